// Vision-4 / Vision-6 hash on the device: the permutation
// (verifier/verifier/src/hash/vision_{4,6}/permutation.rs), the sponge digest
// (digest.rs), the Vision-4 compression (vision_4/compression.rs) and Merkle
// trees built from them (BinaryMerkleTreeProver with VisionHasherDigest leaves,
// prover/examples/src/lib.rs setup_vision4, prover/prover/benches/binary_merkle_tree.rs).
//
// Only the hash's defining data is held here: the coefficients of the affine
// map B(x) = c0 + c1 x + c2 x^2 + c3 x^4 and the inputs of the round-key
// schedule.  The round constants, B^-1 and every lookup table are derived on
// the host when a device first uses a width:
//  * round constants: the key schedule of the reference's constants tests, run
//    through this file's own S-box and MDS;
//  * B^-1: B's linear part inverted over GF(2) (Gauss-Jordan on its 128 x 128
//    bit matrix), plus the constant B^-1(0) = L^-1(c0);
//  * window tables (the 4-bit layout of gf128_dev.hpp's win_mul_perm) for the
//    F2-linear maps the S-box needs: x^(2^64) and the linear parts of B and
//    B^-1 (32 rows each), the Frobenius maps of the subfield chain below on the
//    words that determine a subfield element (16, 8 and 8 rows), and the
//    inverses of GF(2^8)^*.
//
// On the device one lane owns one state.  The S-box's batch inversion
// (verifier/math/src/batch_invert.rs: invert_or_zero of every element) uses one
// field inversion per S-box for the whole state (prefix products, then back
// substitution).  The inversion goes down the subfields by norms:
// n1 = a a^(2^64) lies in GF(2^64), n2 = n1 n1^(2^32) in GF(2^32),
// n3 = n2 n2^(2^16) in GF(2^16), n4 = n3 n3^(2^8) in GF(2^8), whose inverse is a
// table entry; then a^-1 = a^(2^64) n1^(2^32) n2^(2^16) n3^(2^8) n4^-1.  That is
// 8 general products and 57 table reads (the Itoh-Tsujii chain a^(2^128-2) takes
// 10 products and 12 full maps, 384 reads).  In the GHASH representation an
// element of GF(2^64) is fixed by its words 0 and 2, one of GF(2^32) or GF(2^16)
// by word 0 and one of GF(2^8) by nibbles 0 and 4, so the subfield maps read
// only those nibbles.  The tables (36 KiB) and the round constants sit in LDS
// for the whole launch.
#include <hip/hip_runtime.h>
#include <string.h>

#include <mutex>
#include <new>
#include <utility>

#include "abi_util.hpp"
#include "gf128_dev.hpp"
#include "internal.hpp"

using b64::abi::as_stream;
using b64::host::u128;

namespace b64 {
namespace vision {
namespace {

struct u64x2 {
    uint64_t lo, hi;
};

// B(x) = c0 + c1 x + c2 x^2 + c3 x^4 (constants.rs B_FWD_COEFFS, both widths)
constexpr u64x2 B_FWD[4] = {
    {0x4c29e4d892ca33ddull, 0xb0b849b207a0f1c7ull}, {0x855eeb8ce24c9523ull, 0xf2240891aac3c3a5ull},
    {0x9d79908388dd768dull, 0xb017e96797ef1fbeull}, {0x00dbb86a2cb472bbull, 0x2bf9f0b94c2b2cebull},
};
// Vision-4 round-key schedule inputs (vision_4/constants.rs)
constexpr u64x2 INITIAL4[4] = {
    {0xf4a1d28234ffdf87ull, 0xd0e015190f5e0795ull}, {0xb93ec3a23ec7681bull, 0x6111731acd9a89f6ull},
    {0x18a34a96728b8d29ull, 0x15da8de707ee3f39ull}, {0xed2a216b1d232bfbull, 0x2ea920e89fbb13a1ull},
};
constexpr u64x2 MATRIX4[16] = {
    {0x8ef96344bc8cbd0full, 0x1cac9af051191d2cull}, {0x4750ab0a854c3a4dull, 0x3ceb350e8c2d2f4dull},
    {0x8e092e899156946bull, 0x09cc78f7eafef94dull}, {0xa3d58f7eedb30dbbull, 0x1d1ad17c8c7ee715ull},
    {0xdd0cdae9dff5dfc5ull, 0x06bee572113950b3ull}, {0x8b4bfa1a5cd3d295ull, 0x4799c3743a356042ull},
    {0x4cb89ac6bf636d0bull, 0x516883bf97b07fcfull}, {0xdf1a9113b04d755bull, 0x0b681d1621d6aa1full},
    {0x3ddfe74c125e99e5ull, 0xe9ba2f01051f6b4bull}, {0x33e31cfb26b916d1ull, 0x1f58efc5ecf2e1b9ull},
    {0xf67efc573241fe19ull, 0x36da671249b5444cull}, {0x615a130556e5eb99ull, 0x056cae18b867d486ull},
    {0x8dd9df84956bcf11ull, 0x1d8645dd4b1e46b7ull}, {0x95efe6255339375bull, 0x3efa34e48e3b2183ull},
    {0x97d0812db56c5effull, 0x79bf44bae0d23793ull}, {0x9e538151daddae85ull, 0x7fb3b7fce84776e3ull},
};
constexpr u64x2 CONSTANT4[4] = {
    {0x93ca17e6bb10203full, 0x0d42981a71a7c2e4ull}, {0xca17ed48e18717b9ull, 0x9ceed672ccb7030full},
    {0x95c6ecdb4a349d15ull, 0x0cec5c463024f2baull}, {0x990418f3a23e3c99ull, 0xa46023576301b995ull},
};
// Vision-6 round-key schedule inputs (vision_6/constants.rs)
constexpr u64x2 INITIAL6[6] = {
    {0xf4a1d28234ffdf87ull, 0xd0e015190f5e0795ull}, {0xb93ec3a23ec7681bull, 0x6111731acd9a89f6ull},
    {0x18a34a96728b8d29ull, 0x15da8de707ee3f39ull}, {0xed2a216b1d232bfbull, 0x2ea920e89fbb13a1ull},
    {0x8ef96344bc8cbd0full, 0x1cac9af051191d2cull}, {0x4750ab0a854c3a4dull, 0x3ceb350e8c2d2f4dull},
};
constexpr u64x2 MATRIX6[36] = {
    {0x8e092e899156946bull, 0x09cc78f7eafef94dull}, {0xa3d58f7eedb30dbbull, 0x1d1ad17c8c7ee715ull},
    {0xdd0cdae9dff5dfc5ull, 0x06bee572113950b3ull}, {0x8b4bfa1a5cd3d295ull, 0x4799c3743a356042ull},
    {0x4cb89ac6bf636d0bull, 0x516883bf97b07fcfull}, {0xdf1a9113b04d755bull, 0x0b681d1621d6aa1full},
    {0x3ddfe74c125e99e5ull, 0xe9ba2f01051f6b4bull}, {0x33e31cfb26b916d1ull, 0x1f58efc5ecf2e1b9ull},
    {0xf67efc573241fe19ull, 0x36da671249b5444cull}, {0x615a130556e5eb99ull, 0x056cae18b867d486ull},
    {0x8dd9df84956bcf11ull, 0x1d8645dd4b1e46b7ull}, {0x95efe6255339375bull, 0x3efa34e48e3b2183ull},
    {0x97d0812db56c5effull, 0x79bf44bae0d23793ull}, {0x9e538151daddae85ull, 0x7fb3b7fce84776e3ull},
    {0x93ca17e6bb10203full, 0x0d42981a71a7c2e4ull}, {0xca17ed48e18717b9ull, 0x9ceed672ccb7030full},
    {0x95c6ecdb4a349d15ull, 0x0cec5c463024f2baull}, {0x990418f3a23e3c99ull, 0xa46023576301b995ull},
    {0x5d33ae5d97eb83afull, 0x6227628fea1dc546ull}, {0x6d94bf331473f6f1ull, 0x2ef550b7cda1e285ull},
    {0x698bc6ae2e3a8499ull, 0x375f4232fa7ca41aull}, {0x012b9ab22c303529ull, 0x2d2f859caba39ed7ull},
    {0x0d0bf7b6399257b7ull, 0x5672a0afa29cac30ull}, {0xf3ae1aee7aa41593ull, 0xebd30b6903f1fe11ull},
    {0x6f66f0ef2f18bd8dull, 0x39c54aa0738c373aull}, {0xb89afcd0d6a1a9f9ull, 0x083a52acda7952eaull},
    {0xbafc31f177df4359ull, 0x07c782c6beef83baull}, {0xe5ca6278e76adc93ull, 0x1ed204d4affa05b8ull},
    {0xa543ee1ec99323cbull, 0x5ce4a5c3be5e0bc5ull}, {0xe5f49f51af3c2df7ull, 0x0d54d1bf8ea15dafull},
    {0xdc4e99bc400b9fa9ull, 0x22510fa8fffa288aull}, {0xf8135f4f4b2b4491ull, 0x78ca292f525f6a68ull},
    {0x79afa9e365bfa159ull, 0x0ef5f99b1b4a36e9ull}, {0x3a2f3f77630cffa1ull, 0x7a999d6e5f0b133dull},
    {0xd21d9ad73aef8469ull, 0xb1ae2573d4a15bb6ull}, {0x2535d2146c4869e1ull, 0x9c348d3213f5a078ull},
};
constexpr u64x2 CONSTANT6[6] = {
    {0xb4ee8dd307ef621full, 0xabd5e3e6c590952dull}, {0x2cf0fc6f51bedcebull, 0x22681ba7c4bf168cull},
    {0x1a74f42c784257a1ull, 0x41baa8130ac4ddf6ull}, {0x11ac77cfb11d7ccdull, 0x26f2ab43dac65cb0ull},
    {0x1552b09907e977e3ull, 0x29218b482b8935faull}, {0xfe818a702d07fb9dull, 0x001cb764573d0acdull},
};

constexpr int ROUNDS = 8;
constexpr int RC_ROWS = 1 + 2 * ROUNDS;
constexpr int MAXM = 6;
// full 32-row window tables; T_P64: x -> x^(2^64)
constexpr int N_TABLES = 3;
enum Table { T_P64 = 0, T_BFWD, T_BINV };
// Partial tables after them (uint4 offsets, each 256-byte aligned): rows for
// the nibbles that fix a subfield element, entry n of row r at slot n ^ (r & 15).
constexpr int PT_F32 = N_TABLES * WIN_ENTRIES;  // x^(2^32) on GF(2^64): rows 0..7 word 0, 8..15 word 2
constexpr int PT_F16 = PT_F32 + 16 * 16;        // x^(2^16) on GF(2^32): word 0
constexpr int PT_F8 = PT_F16 + 8 * 16;          // x^(2^8) on GF(2^16): word 0
constexpr int PT_I8 = PT_F8 + 8 * 16;           // x^-1 on GF(2^8)^*, by nibble 0 | nibble 4 << 4 of word 0
constexpr int VBLOCK = 256;
// Two blocks of VBLOCK threads share a CU (37 KiB of LDS each), so kernels are
// held to two waves per SIMD: at most 256 VGPRs and AGPRs together.
#define VIS_KERNEL __global__ __launch_bounds__(VBLOCK) __attribute__((amdgpu_waves_per_eu(2)))
// device data of one width: tables, then round constants (RC_ROWS * MAXM), then c0 and B^-1(0)
constexpr int DEV_TAB = PT_I8 + 256;
constexpr int DEV_RC = DEV_TAB;
constexpr int DEV_C = DEV_RC + RC_ROWS * MAXM;
constexpr int DEV_ELEMS = DEV_C + 2;

inline u128 to128(u64x2 v) { return ((u128)v.hi << 64) | v.lo; }
inline uint4 u4(u128 v) {
    return make_uint4((uint32_t)v, (uint32_t)(v >> 32), (uint32_t)(v >> 64), (uint32_t)(v >> 96));
}

// ---- host side: the derived data -------------------------------------------
struct HostVision {
    int m = 0;
    u128 c[4];              // B's coefficients
    u128 binv_cols[128];    // L^-1(x^j)
    u128 binv0 = 0;         // B^-1(0)
    u128 rc[RC_ROWS][MAXM];

    u128 sqr(u128 x) const { return host::gf_mul(x, x); }
    u128 b_fwd_linear(u128 x) const {
        const u128 x2 = sqr(x), x4 = sqr(x2);
        return host::gf_mul(c[1], x) ^ host::gf_mul(c[2], x2) ^ host::gf_mul(c[3], x4);
    }
    u128 b_inv_linear(u128 y) const {
        u128 acc = 0;
        for (int j = 0; j < 128; j++)
            if ((y >> j) & 1) acc ^= binv_cols[j];
        return acc;
    }
    void sbox(u128 *s, bool inverse) const {
        for (int i = 0; i < m; i++) {
            const u128 x = s[i] ? host::gf_inv(s[i]) : 0;
            s[i] = inverse ? (b_inv_linear(x) ^ binv0) : (b_fwd_linear(x) ^ c[0]);
        }
    }
    void mds(u128 *a) const {
        u128 o[MAXM];
        if (m == 4) {
            // rows [2 3 1 1], [1 2 3 1], [1 1 2 3], [3 1 1 2] with 2 = x, 3 = x + 1
            for (int i = 0; i < 4; i++)
                o[i] = host::gf_mul(a[i] ^ a[(i + 1) & 3], 2) ^ a[(i + 1) & 3] ^ a[(i + 2) & 3] ^ a[(i + 3) & 3];
        } else {
            // circulant, first row [x^-1, x^-1, x, 1, x, x + 1]
            const u128 inv_x = host::gf_inv(2);
            for (int i = 0; i < 6; i++) {
                const u128 *r = a;
                auto at = [&](int k) { return r[(i + k) % 6]; };
                o[i] = host::gf_mul(at(0) ^ at(1), inv_x) ^ host::gf_mul(at(2) ^ at(4) ^ at(5), 2) ^ at(3) ^ at(5);
            }
        }
        memcpy(a, o, sizeof(u128) * (size_t)m);
    }

    // Gauss-Jordan over GF(2) on the columns L(x^j) of B's linear part.
    bool invert_b() {
        u128 rows[128], inv[128];
        for (int r = 0; r < 128; r++) {
            rows[r] = 0;
            inv[r] = (u128)1 << r;
        }
        for (int j = 0; j < 128; j++) {
            const u128 col = b_fwd_linear((u128)1 << j);
            for (int r = 0; r < 128; r++)
                if ((col >> r) & 1) rows[r] |= (u128)1 << j;
        }
        for (int c0 = 0; c0 < 128; c0++) {
            int p = c0;
            while (p < 128 && !((rows[p] >> c0) & 1)) p++;
            if (p == 128) return false;
            std::swap(rows[p], rows[c0]);
            std::swap(inv[p], inv[c0]);
            for (int r = 0; r < 128; r++)
                if (r != c0 && ((rows[r] >> c0) & 1)) {
                    rows[r] ^= rows[c0];
                    inv[r] ^= inv[c0];
                }
        }
        for (int j = 0; j < 128; j++) {
            u128 col = 0;
            for (int r = 0; r < 128; r++)
                if ((inv[r] >> j) & 1) col |= (u128)1 << r;
            binv_cols[j] = col;
        }
        binv0 = b_inv_linear(c[0]);
        return true;
    }

    // The round-key schedule: key_injection <- CONSTANTS_MATRIX * key_injection
    // + CONSTANTS_CONSTANT, key_state <- MDS(S-box(key_state)) + key_injection,
    // alternating the S-box's B^-1 and B halves.
    void round_constants(const u64x2 *initial, const u64x2 *matrix, const u64x2 *constant) {
        u128 state[MAXM], inj[MAXM];
        for (int i = 0; i < m; i++) state[i] = inj[i] = rc[0][i] = to128(initial[i]);
        int idx = 1;
        for (int r = 0; r < ROUNDS; r++)
            for (int half = 0; half < 2; half++) {
                u128 t[MAXM];
                for (int i = 0; i < m; i++) {
                    t[i] = to128(constant[i]);
                    for (int j = 0; j < m; j++) t[i] ^= host::gf_mul(to128(matrix[i * m + j]), inj[j]);
                }
                memcpy(inj, t, sizeof(u128) * (size_t)m);
                sbox(state, half == 0);
                mds(state);
                for (int i = 0; i < m; i++) state[i] ^= inj[i];
                memcpy(rc[idx++], state, sizeof(u128) * (size_t)m);
            }
    }
};

// Window table of a linear map given by its column images col[j] = L(e_j) for
// the input bits e_j it reads: row k, entry n = L(n e_(4k)), stored at slot
// n ^ (k & 15) of the row.
void window_table(const u128 *col, uint4 *tab, int rows = 32) {
    for (int k = 0; k < rows; k++)
        for (int n = 0; n < 16; n++) {
            u128 v = 0;
            for (int b = 0; b < 4; b++)
                if ((n >> b) & 1) v ^= col[4 * k + b];
            tab[k * 16 + (n ^ (k & 15))] = u4(v);
        }
}

// ---- the subfield chain of the S-box inversion --------------------------------
// F2-linear maps as their 128 column images.
struct Lin {
    u128 col[128];
    u128 operator()(u128 x) const {
        u128 r = 0;
        for (int j = 0; j < 128; j++)
            if ((x >> j) & 1) r ^= col[j];
        return r;
    }
    Lin then(const Lin &g) const {  // g after this
        Lin r;
        for (int j = 0; j < 128; j++) r.col[j] = g(col[j]);
        return r;
    }
};

// A basis of GF(2^d) inside GF(2^128): the kernel of x -> x^(2^d) + x (Gauss
// elimination over F2, tracking which input bits sum to each reduced vector).
int subfield_basis(const Lin &frob_d, u128 *basis) {
    u128 piv_v[128], piv_c[128];
    bool has[128] = {};
    int n = 0;
    for (int j = 0; j < 128; j++) {
        u128 v = frob_d.col[j] ^ ((u128)1 << j), c = (u128)1 << j;
        while (v) {
            int h = 127;
            while (!((v >> h) & 1)) h--;
            if (!has[h]) {
                has[h] = true;
                piv_v[h] = v;
                piv_c[h] = c;
                break;
            }
            v ^= piv_v[h];
            c ^= piv_c[h];
        }
        if (!v) basis[n++] = c;
    }
    return n;
}

// Given a basis of a d-dimensional subspace V and a projection p onto d bits,
// lift[j] = the element of V whose projection is bit j alone.  False when p is
// not one-to-one on V.
template <class Proj>
bool lift_basis(const u128 *basis, int d, Proj p, u128 *lift) {
    u128 pv[128], el[128];
    for (int i = 0; i < d; i++) {
        pv[i] = p(basis[i]);
        el[i] = basis[i];
    }
    for (int j = 0; j < d; j++) {
        int r = j;
        while (r < d && !((pv[r] >> j) & 1)) r++;
        if (r == d) return false;
        std::swap(pv[r], pv[j]);
        std::swap(el[r], el[j]);
        for (int i = 0; i < d; i++)
            if (i != j && ((pv[i] >> j) & 1)) {
                pv[i] ^= pv[j];
                el[i] ^= el[j];
            }
    }
    for (int j = 0; j < d; j++) lift[j] = el[j];
    return true;
}

inline u128 proj_w02(u128 x) { return (x & 0xFFFFFFFFu) | (((x >> 64) & 0xFFFFFFFFu) << 32); }
inline u128 proj_w0(u128 x) { return x & 0xFFFFFFFFu; }
inline uint32_t idx_i8(uint32_t w0) { return (w0 & 0xFu) | ((w0 >> 12) & 0xF0u); }
inline u128 proj_i8(u128 x) { return idx_i8((uint32_t)x); }

inline u128 from4(uint4 v) {
    return ((u128)v.w << 96) | ((u128)v.z << 64) | ((u128)v.y << 32) | v.x;
}
// What the device's lookups compute: rows row0.. of a table on the nibbles of w.
inline u128 rows_apply(const uint4 *tab, int row0, uint32_t w, int rows = 8) {
    u128 r = 0;
    for (int k = 0; k < rows; k++) {
        const int row = row0 + k;
        r ^= from4(tab[row * 16 + (((w >> (4 * k)) & 15u) ^ (uint32_t)(row & 15))]);
    }
    return r;
}

// The inversion's tables: T_P64 and PT_F32 .. PT_I8 of the layout above, then a
// check of the chain, exactly as the device runs it, on a few elements.
bool inversion_tables(uint4 *out) {
    Lin f8;
    for (int j = 0; j < 128; j++) {
        u128 x = (u128)1 << j;
        for (int k = 0; k < 8; k++) x = host::gf_mul(x, x);
        f8.col[j] = x;
    }
    const Lin f16 = f8.then(f8), f32 = f16.then(f16), f64 = f32.then(f32);
    window_table(f64.col, out + T_P64 * WIN_ENTRIES);
    u128 b64[128], b32[128], b8[128], l64[64], l32[32], l8[8];
    if (subfield_basis(f64, b64) != 64 || subfield_basis(f32, b32) != 32 || subfield_basis(f8, b8) != 8)
        return false;
    if (!lift_basis(b64, 64, proj_w02, l64) || !lift_basis(b32, 32, proj_w0, l32) ||
        !lift_basis(b8, 8, proj_i8, l8))
        return false;
    u128 col[64];
    for (int j = 0; j < 64; j++) col[j] = f32(l64[j]);
    window_table(col, out + PT_F32, 16);
    for (int j = 0; j < 32; j++) col[j] = f16(l32[j]);
    window_table(col, out + PT_F16, 8);
    // GF(2^16) lies in GF(2^32), so word 0 fixes its elements as well
    for (int j = 0; j < 32; j++) col[j] = f8(l32[j]);
    window_table(col, out + PT_F8, 8);
    for (uint32_t i = 0; i < 256; i++) {
        u128 v = 0;
        for (int j = 0; j < 8; j++)
            if ((i >> j) & 1) v ^= l8[j];
        // v^254 = v^-1 on GF(2^8)^*
        u128 inv = 1, sq = v;
        for (int e = 254; e; e >>= 1) {
            if (e & 1) inv = host::gf_mul(inv, sq);
            sq = host::gf_mul(sq, sq);
        }
        out[PT_I8 + i] = u4(i ? inv : 0);
    }
    u128 a = ((u128)0x9e3779b97f4a7c15ull << 64) | 0xbf58476d1ce4e5b9ull;
    for (int t = 0; t < 8; t++, a = host::gf_mul(a, a) ^ t) {
        const u128 t64 = rows_apply(out + T_P64 * WIN_ENTRIES, 0, (uint32_t)a) ^
                         rows_apply(out + T_P64 * WIN_ENTRIES, 8, (uint32_t)(a >> 32)) ^
                         rows_apply(out + T_P64 * WIN_ENTRIES, 16, (uint32_t)(a >> 64)) ^
                         rows_apply(out + T_P64 * WIN_ENTRIES, 24, (uint32_t)(a >> 96));
        const u128 n1 = host::gf_mul(a, t64);
        const u128 t32 = rows_apply(out + PT_F32, 0, (uint32_t)n1) ^ rows_apply(out + PT_F32, 8, (uint32_t)(n1 >> 64));
        const u128 n2 = host::gf_mul(n1, t32);
        const u128 t16 = rows_apply(out + PT_F16, 0, (uint32_t)n2);
        const u128 n3 = host::gf_mul(n2, t16);
        const u128 t8 = rows_apply(out + PT_F8, 0, (uint32_t)n3);
        const u128 n4 = host::gf_mul(n3, t8);
        const u128 i8 = from4(out[PT_I8 + idx_i8((uint32_t)n4)]);
        const u128 inv = host::gf_mul(host::gf_mul(host::gf_mul(host::gf_mul(t64, t32), t16), t8), i8);
        if (a != 0 && host::gf_mul(a, inv) != 1) return false;
    }
    return true;
}

bool build_device_data(int m, uint4 *out) {
    HostVision hv;
    hv.m = m;
    for (int i = 0; i < 4; i++) hv.c[i] = to128(B_FWD[i]);
    if (!hv.invert_b()) return false;
    if (m == 4)
        hv.round_constants(INITIAL4, MATRIX4, CONSTANT4);
    else
        hv.round_constants(INITIAL6, MATRIX6, CONSTANT6);
    if (!inversion_tables(out)) return false;
    u128 col[128];
    for (int j = 0; j < 128; j++) col[j] = hv.b_fwd_linear((u128)1 << j);
    window_table(col, out + T_BFWD * WIN_ENTRIES);
    window_table(hv.binv_cols, out + T_BINV * WIN_ENTRIES);
    for (int r = 0; r < RC_ROWS; r++)
        for (int i = 0; i < MAXM; i++) out[DEV_RC + r * MAXM + i] = u4(i < m ? hv.rc[r][i] : 0);
    out[DEV_C] = u4(hv.c[0]);
    out[DEV_C + 1] = u4(hv.binv0);
    return true;
}

// Per device and width: the derived data, uploaded once.
const uint4 *device_data(int m) {
    static std::mutex mu;
    static const uint4 *cache[64][2] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    const int w = m == 4 ? 0 : 1;
    std::lock_guard<std::mutex> lock(mu);
    if (cache[dev][w]) return cache[dev][w];
    uint4 *host = new (std::nothrow) uint4[DEV_ELEMS];
    if (!host) return nullptr;
    uint4 *d = nullptr;
    const bool ok = build_device_data(m, host) && hipMalloc(&d, DEV_ELEMS * sizeof(uint4)) == hipSuccess &&
                    hipMemcpy(d, host, DEV_ELEMS * sizeof(uint4), hipMemcpyHostToDevice) == hipSuccess;
    delete[] host;
    if (!ok) {
        if (d) (void)hipFree(d);
        return nullptr;
    }
    cache[dev][w] = d;  // held until the process ends
    return d;
}

// ---- device side -------------------------------------------------------------
// The permutation is written once over an "ops" policy that supplies the
// general product and the table maps:
//  * LaneOps: one lane owns a state and does all of its work (throughput form);
//  * QuadOps: the four lanes of a quad (lanes 4i .. 4i+3) hold the same state and
//    split every product and every map four ways, exchanging partial results
//    with DPP quad permutes.  A product costs each lane one 64x64 carry-less
//    product (a_(q>>1) * b_(q&1)) instead of three, a map 8 table reads instead
//    of 32, so a state's dependent chain is about 2.5x shorter (latency form,
//    for launches with too few states to fill the device).
struct LaneOps {
    __device__ __forceinline__ uint4 mul(uint4 a, uint4 b) const { return gf_mul(a, b); }
    __device__ __forceinline__ uint4 map(uint32_t tabs, int t, uint4 v) const {
        return win_mul_perm(tabs + (uint32_t)(t * WIN_ENTRIES * 16), v);
    }
    __device__ __forceinline__ bool writer() const { return true; }
    __device__ __forceinline__ void set_lane(uint32_t) {}
};

// Lane K of the quad, to every lane of the quad.
template <int K>
__device__ __forceinline__ uint32_t quad_bcast(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, K * 0x55, 0xF, 0xF, true);
}
template <int K>
__device__ __forceinline__ uint64_t quad_bcast64(uint64_t v) {
    return ((uint64_t)quad_bcast<K>((uint32_t)(v >> 32)) << 32) | quad_bcast<K>((uint32_t)v);
}
// XOR of the four lanes' values, in every lane (quad_perm [1,0,3,2], then [2,3,0,1]).
__device__ __forceinline__ uint32_t quad_xor(uint32_t v) {
    v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, true);
    return v ^ (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, true);
}

struct QuadOps {
    uint32_t q;  // lane within the quad
    // a b = a0 b0 + (a0 b1 + a1 b0) X + a1 b1 X^2 with X = x^64; lane q computes
    // a_(q>>1) b_(q&1).
    __device__ __forceinline__ uint4 mul(uint4 a, uint4 b) const {
        const bool ah = q & 2, bh = q & 1;
        const uint64_t av = ah ? (((uint64_t)a.w << 32) | a.z) : (((uint64_t)a.y << 32) | a.x);
        const uint64_t bv = bh ? (((uint64_t)b.w << 32) | b.z) : (((uint64_t)b.y << 32) | b.x);
        uint64_t lo, hi;
        clmul64(av, bv, lo, hi);
        const uint64_t p0l = quad_bcast64<0>(lo), p0h = quad_bcast64<0>(hi);
        const uint64_t p1l = quad_bcast64<1>(lo), p1h = quad_bcast64<1>(hi);
        const uint64_t p2l = quad_bcast64<2>(lo), p2h = quad_bcast64<2>(hi);
        const uint64_t p3l = quad_bcast64<3>(lo), p3h = quad_bcast64<3>(hi);
        return gf_reduce(p0l, xor3_64(p0h, p1l, p2l), xor3_64(p1h, p2h, p3l), p3h);
    }
    // Lane q looks up word q of v in rows 8q .. 8q+7 of the table (those rows'
    // slot swizzle is the nibble index XOR 8 (q & 1)), then the quad sums.
    __device__ __forceinline__ uint4 map(uint32_t tabs, int t, uint4 v) const {
        const uint32_t w = q == 0 ? v.x : (q == 1 ? v.y : (q == 2 ? v.z : v.w));
        const uint32_t base = tabs + (uint32_t)(t * WIN_ENTRIES * 16) + q * (8 * 16 * 16);
        const uint4 acc = win_word_perm(base, w ^ ((q & 1) ? 0x88888888u : 0u), 0, make_uint4(0, 0, 0, 0));
        return make_uint4(quad_xor(acc.x), quad_xor(acc.y), quad_xor(acc.z), quad_xor(acc.w));
    }
    __device__ __forceinline__ bool writer() const { return q == 0; }
    __device__ __forceinline__ void set_lane(uint32_t x) { q = x & 3; }
};

// XOR of the 16 lanes of a DPP row, in every lane of the row (quad_perm
// [1,0,3,2], quad_perm [2,3,0,1], row_ror:4, row_ror:8).
__device__ __forceinline__ uint32_t row_xor(uint32_t v) {
    v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, true);
    v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, true);
    v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x124, 0xF, 0xF, true);
    return v ^ (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xF, 0xF, true);
}

__device__ __forceinline__ uint32_t word_of(uint4 v, uint32_t i) {
    return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w));
}

// Sixteen lanes (a DPP row) per state: lane l computes the 32x32 carry-less
// product of word l>>2 of a and word l&3 of b and places it in a 256-bit frame
// at word (l>>2)+(l&3); the row XOR of the frames is the full product.  A map
// takes one byte of v per lane (two table rows).  The DPP reduction chains
// (four dependent steps per word) eat most of the gain: a narrow tree level
// takes 0.146 ms against the quad form's 0.172 ms.
struct RowOps {
    uint32_t l;  // lane within the row
    __device__ __forceinline__ uint4 mul(uint4 a, uint4 b) const {
        const uint32_t i = l >> 2, j = l & 3, at = i + j;
        const uint64_t p = clmul32(word_of(a, i), word_of(b, j));
        const uint32_t lo = (uint32_t)p, hi = (uint32_t)(p >> 32);
        uint32_t f[8];
#pragma unroll
        for (uint32_t k = 0; k < 8; k++) f[k] = row_xor(k == at ? lo : (k == at + 1 ? hi : 0u));
        return gf_reduce(((uint64_t)f[1] << 32) | f[0], ((uint64_t)f[3] << 32) | f[2], ((uint64_t)f[5] << 32) | f[4],
                         ((uint64_t)f[7] << 32) | f[6]);
    }
    // rows 2l and 2l+1 of the table: nibbles 2l and 2l+1 of v; entry n of
    // row r sits at slot n ^ (r & 15)
    __device__ __forceinline__ uint4 map(uint32_t tabs, int t, uint4 v) const {
        const uint32_t byte = (word_of(v, l >> 2) >> (8 * (l & 3))) & 0xFFu;
        const uint32_t r0 = 2 * l, r1 = r0 + 1;
        const uint32_t base = tabs + (uint32_t)(t * WIN_ENTRIES * 16);
        const uint32_t a0 = base + r0 * 256 + (((byte & 15u) ^ (r0 & 15u)) << 4);
        const uint32_t a1 = base + r1 * 256 + (((byte >> 4) ^ (r1 & 15u)) << 4);
        const u32x4 e0 = *(const lds_cu32x4 *)(uintptr_t)a0;
        const u32x4 e1 = *(const lds_cu32x4 *)(uintptr_t)a1;
        return make_uint4(row_xor(e0.x ^ e1.x), row_xor(e0.y ^ e1.y), row_xor(e0.z ^ e1.z), row_xor(e0.w ^ e1.w));
    }
    __device__ __forceinline__ bool writer() const { return l == 0; }
    __device__ __forceinline__ void set_lane(uint32_t x) { l = x & 15; }
};

template <int LANES>
struct OpsFor;
template <>
struct OpsFor<1> {
    using type = LaneOps;
};
template <>
struct OpsFor<4> {
    using type = QuadOps;
};
template <>
struct OpsFor<16> {
    using type = RowOps;
};
constexpr int lanes_log(int lanes) { return lanes == 16 ? 4 : (lanes == 4 ? 2 : 0); }

// v * x^-1 (ghash.rs mul_inv_x)
__device__ __forceinline__ uint4 gf_mul_inv_x(uint4 v) {
    const uint32_t m = 0u - (v.x & 1u);
    return make_uint4(((v.x >> 1) | (v.y << 31)) ^ (m & 0x43u), (v.y >> 1) | (v.z << 31), (v.z >> 1) | (v.w << 31),
                      (v.w >> 1) ^ (m & 0x80000000u));
}

// a^-1 for a != 0 down the subfield chain (norms into GF(2^64), GF(2^32),
// GF(2^16), GF(2^8), a table inverse there, and the Frobenius images multiplied
// back up).  The subfield maps and the GF(2^8) entry are read by every lane of a
// state alike; only the full map and the products go through the ops policy.
template <class Ops>
__device__ uint4 vis_inv(const Ops &o, uint32_t tabs, uint4 a) {
    const uint4 z = make_uint4(0, 0, 0, 0);
    const uint32_t f32 = tabs + PT_F32 * 16, f16 = tabs + PT_F16 * 16, f8 = tabs + PT_F8 * 16;
    const uint4 t64 = o.map(tabs, T_P64, a);
    const uint4 n1 = o.mul(a, t64);
    const uint4 t32 = win_word_perm(f32, n1.z, 8, win_word_perm(f32, n1.x, 0, z));
    const uint4 n2 = o.mul(n1, t32);
    const uint4 p32 = o.mul(t64, t32);
    const uint4 t16 = win_word_perm(f16, n2.x, 0, z);
    const uint4 n3 = o.mul(n2, t16);
    const uint4 p16 = o.mul(p32, t16);
    const uint4 t8 = win_word_perm(f8, n3.x, 0, z);
    const uint4 n4 = o.mul(n3, t8);
    const uint4 p8 = o.mul(p16, t8);
    const uint32_t i8 = tabs + (uint32_t)(PT_I8 * 16) + (((n4.x << 4) & 0xF0u) | ((n4.x >> 8) & 0xF00u));
    const u32x4 e = *(const lds_cu32x4 *)(uintptr_t)i8;
    return o.mul(p8, make_uint4(e.x, e.y, e.z, e.w));
}

// invert_or_zero of every element with one inversion (zeros enter as ones and
// come back as zeros).
template <int M, class Ops>
__device__ __forceinline__ void vis_batch_inv(const Ops &o, uint32_t tabs, uint4 (&s)[M]) {
    const uint4 one = make_uint4(1, 0, 0, 0), zero = make_uint4(0, 0, 0, 0);
    uint4 pre[M], x[M];
    bool z[M];
    uint4 acc = one;
#pragma unroll
    for (int i = 0; i < M; i++) {
        z[i] = gzero(s[i]);
        x[i] = z[i] ? one : s[i];
        pre[i] = acc;
        acc = i == 0 ? x[0] : o.mul(acc, x[i]);
    }
    uint4 inv = vis_inv(o, tabs, acc);
#pragma unroll
    for (int i = M - 1; i >= 0; i--) {
        const uint4 out = i == 0 ? inv : o.mul(inv, pre[i]);
        if (i > 0) inv = o.mul(inv, x[i]);
        s[i] = z[i] ? zero : out;
    }
}

template <int M>
__device__ __forceinline__ void vis_mds(uint4 (&a)[M]) {
    if constexpr (M == 4) {
        const uint4 sum = gxor(gxor(a[0], a[1]), gxor(a[2], a[3]));
        const uint4 a0 = a[0];
        a[0] = gxor(gxor(a[0], sum), gf_mul_x(gxor(a[0], a[1])));
        a[1] = gxor(gxor(a[1], sum), gf_mul_x(gxor(a[1], a[2])));
        a[2] = gxor(gxor(a[2], sum), gf_mul_x(gxor(a[2], a[3])));
        a[3] = gxor(gxor(a[3], sum), gf_mul_x(gxor(a[3], a0)));
    } else {
        uint4 o[6];
#pragma unroll
        for (int i = 0; i < 6; i++) {
            const uint4 p = gf_mul_inv_x(gxor(a[i], a[(i + 1) % 6]));
            const uint4 q = gf_mul_x(gxor(gxor(a[(i + 2) % 6], a[(i + 4) % 6]), a[(i + 5) % 6]));
            o[i] = gxor(gxor(p, q), gxor(a[(i + 3) % 6], a[(i + 5) % 6]));
        }
#pragma unroll
        for (int i = 0; i < 6; i++) a[i] = o[i];
    }
}

// LDS view of one width's data.
struct VisionLds {
    uint32_t tabs;      // LDS address of the 8 tables
    const uint4 *rc;    // RC_ROWS * MAXM
    uint4 c0, binv0;
};

template <int M, class Ops = LaneOps>
__device__ void vis_permute(const VisionLds &v, uint4 (&s)[M], const Ops &o = Ops()) {
#pragma unroll
    for (int i = 0; i < M; i++) s[i] = gxor(s[i], v.rc[i]);
#pragma unroll 1
    for (int r = 0; r < ROUNDS; r++) {
        vis_batch_inv<M>(o, v.tabs, s);
#pragma unroll
        for (int i = 0; i < M; i++) s[i] = gxor(o.map(v.tabs, T_BINV, s[i]), v.binv0);
        vis_mds<M>(s);
#pragma unroll
        for (int i = 0; i < M; i++) s[i] = gxor(s[i], v.rc[(1 + 2 * r) * MAXM + i]);
        vis_batch_inv<M>(o, v.tabs, s);
#pragma unroll
        for (int i = 0; i < M; i++) s[i] = gxor(o.map(v.tabs, T_BFWD, s[i]), v.c0);
        vis_mds<M>(s);
#pragma unroll
        for (int i = 0; i < M; i++) s[i] = gxor(s[i], v.rc[(2 + 2 * r) * MAXM + i]);
    }
}

// Stages one width's tables and constants into LDS (37 KiB).
__device__ VisionLds vis_stage(const uint4 *__restrict__ dev, uint4 *lds) {
    for (int e = threadIdx.x; e < DEV_ELEMS; e += VBLOCK) lds[e] = dev[e];
    __syncthreads();
    return VisionLds{lds_addr(lds), lds + DEV_RC, lds[DEV_C], lds[DEV_C + 1]};
}

template <int M>
VIS_KERNEL void k_vision_permute(const uint4 *__restrict__ dev, uint4 *__restrict__ states,
                                                         uint64_t n) {
    __shared__ __attribute__((aligned(256))) uint4 vis_lds[DEV_ELEMS];
    const VisionLds v = vis_stage(dev, vis_lds);
    for (uint64_t k = (uint64_t)blockIdx.x * VBLOCK + threadIdx.x; k < n; k += (uint64_t)gridDim.x * VBLOCK) {
        uint4 s[M];
#pragma unroll
        for (int i = 0; i < M; i++) s[i] = states[k * M + i];
        vis_permute<M>(v, s);
#pragma unroll
        for (int i = 0; i < M; i++) states[k * M + i] = s[i];
    }
}

// VisionHasherDigest of n messages of len elements each (msg k at
// msgs + k * len): rate R = M - 2 elements; each full block overwrites
// state[0 .. R) and permutes; the final block holds the remaining elements, a
// 0x80 byte after them and 0x01 in its last byte (a whole padding block when
// nothing remains); the digest is state[0] || state[1].
// LANES lanes per state (1, 4 or 16: LaneOps, QuadOps, RowOps); the loop bound
// is uniform over a state's lanes.
template <int M, int LANES>
VIS_KERNEL void k_vision_digest(const uint4 *__restrict__ dev, const uint4 *__restrict__ msgs, uint64_t n,
                                uint32_t len, uint4 *__restrict__ out) {
    constexpr int R = M - 2;
    constexpr int SHIFT = lanes_log(LANES);
    __shared__ __attribute__((aligned(256))) uint4 vis_lds[DEV_ELEMS];
    const VisionLds v = vis_stage(dev, vis_lds);
    typename OpsFor<LANES>::type ops;
    ops.set_lane(threadIdx.x);
    for (uint64_t t = (uint64_t)blockIdx.x * VBLOCK + threadIdx.x; (t >> SHIFT) < n;
         t += (uint64_t)gridDim.x * VBLOCK) {
        const uint64_t k = t >> SHIFT;
        const uint4 *p = msgs + k * len;
        uint4 s[M];
#pragma unroll
        for (int i = 0; i < M; i++) s[i] = make_uint4(0, 0, 0, 0);
        uint32_t e = 0;
        for (; e + R <= len; e += R) {
#pragma unroll
            for (int i = 0; i < R; i++) s[i] = p[e + i];
            vis_permute<M>(v, s, ops);
        }
        const uint32_t rem = len - e;
#pragma unroll
        for (int i = 0; i < R; i++) {
            s[i] = (uint32_t)i < rem ? p[e + i] : make_uint4(0, 0, 0, 0);
            if ((uint32_t)i == rem) s[i].x ^= 0x80u;
        }
        s[R - 1].w ^= 0x01000000u;
        vis_permute<M>(v, s, ops);
        if (ops.writer()) {
            out[2 * k] = s[0];
            out[2 * k + 1] = s[1];
        }
    }
}

// VisionCompression of one tree level: node i of the level above is
// compress(in[2i], in[2i+1]) = first half of permute(l0, l1, r0, r1) + (l0, l1).
// LANES lanes per node, as in k_vision_digest.
template <int LANES>
VIS_KERNEL void k_vision_compress(const uint4 *__restrict__ dev, uint4 *__restrict__ tree, uint64_t in_off,
                                  uint64_t n_out) {
    constexpr int SHIFT = lanes_log(LANES);
    __shared__ __attribute__((aligned(256))) uint4 vis_lds[DEV_ELEMS];
    const VisionLds v = vis_stage(dev, vis_lds);
    typename OpsFor<LANES>::type ops;
    ops.set_lane(threadIdx.x);
    for (uint64_t t = (uint64_t)blockIdx.x * VBLOCK + threadIdx.x; (t >> SHIFT) < n_out;
         t += (uint64_t)gridDim.x * VBLOCK) {
        const uint64_t i = t >> SHIFT;
        const uint4 *in = tree + 2 * (in_off + 2 * i);
        uint4 s[4] = {in[0], in[1], in[2], in[3]};
        const uint4 l0 = s[0], l1 = s[1];
        vis_permute<4>(v, s, ops);
        if (ops.writer()) {
            uint4 *o = tree + 2 * (in_off + 2 * n_out + i);
            o[0] = gxor(s[0], l0);
            o[1] = gxor(s[1], l1);
        }
    }
}

unsigned vis_grid(uint64_t n) {
    const uint64_t b = (n + VBLOCK - 1) / VBLOCK;
    return (unsigned)(b < 4096 ? (b ? b : 1) : 4096);
}

hipError_t launch_permute(int m, const uint4 *dev, uint4 *states, uint64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (m == 4) {
        k_vision_permute<4><<<vis_grid(n), VBLOCK, 0, s>>>(dev, states, n);
    } else {
        k_vision_permute<6><<<vis_grid(n), VBLOCK, 0, s>>>(dev, states, n);
    }
    return hipGetLastError();
}

// Launches with at most this many states run four lanes per state (QuadOps).
// A state's permutations form one dependent chain; with fewer states than
// lanes on the device the quad form's 2.6x shorter chain wins, with more the
// lane form's 1.8x smaller instruction count does (measured: 2^16 Vision-4
// compressions per launch 0.60 ms quad vs 0.52 ms lane, 2^15: 0.31 vs 0.49 ms;
// Vision-6 digests of 2^17 16-element messages 6.3 ms quad vs 5.2 ms lane).
constexpr uint64_t QUAD_MAX_STATES = 1ull << 15;
// ... and with at most this many, sixteen lanes per state (RowOps); measured
// best of 2^10, 2^12 and 2^14 for a 2^17-leaf commit.
constexpr uint64_t ROW_MAX_STATES = 1ull << 12;

int lanes_for(uint64_t n) { return n <= ROW_MAX_STATES ? 16 : (n <= QUAD_MAX_STATES ? 4 : 1); }

template <int M, int LANES>
void launch_digest_l(const uint4 *dev, const uint4 *msgs, uint64_t n, uint32_t len, uint4 *out, hipStream_t s) {
    k_vision_digest<M, LANES><<<vis_grid(LANES * n), VBLOCK, 0, s>>>(dev, msgs, n, len, out);
}

template <int M>
void launch_digest_m(const uint4 *dev, const uint4 *msgs, uint64_t n, uint32_t len, uint4 *out, hipStream_t s) {
    switch (lanes_for(n)) {
        case 16: launch_digest_l<M, 16>(dev, msgs, n, len, out, s); break;
        case 4: launch_digest_l<M, 4>(dev, msgs, n, len, out, s); break;
        default: launch_digest_l<M, 1>(dev, msgs, n, len, out, s); break;
    }
}

hipError_t launch_digest(int m, const uint4 *dev, const uint4 *msgs, uint64_t n, uint32_t len, uint4 *out,
                         hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (m == 4)
        launch_digest_m<4>(dev, msgs, n, len, out, s);
    else
        launch_digest_m<6>(dev, msgs, n, len, out, s);
    return hipGetLastError();
}

hipError_t launch_tree_levels(const uint4 *dev4, uint4 *tree, uint64_t n_leaves, hipStream_t s) {
    uint64_t in_off = 0;
    for (uint64_t n = n_leaves; n > 1; n /= 2) {
        const uint64_t nodes = n / 2;
        switch (lanes_for(nodes)) {
            case 16: k_vision_compress<16><<<vis_grid(16 * nodes), VBLOCK, 0, s>>>(dev4, tree, in_off, nodes); break;
            case 4: k_vision_compress<4><<<vis_grid(4 * nodes), VBLOCK, 0, s>>>(dev4, tree, in_off, nodes); break;
            default: k_vision_compress<1><<<vis_grid(nodes), VBLOCK, 0, s>>>(dev4, tree, in_off, nodes); break;
        }
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        in_off += n;
    }
    return hipSuccess;
}

int check_device() {
    int n = 0;
    return (hipGetDeviceCount(&n) == hipSuccess && n > 0) ? B64_OK : B64_ERR_NO_DEVICE;
}

}  // namespace
}  // namespace vision
}  // namespace b64

namespace b64 {

int vision_merkle_commit(int leaf_m, const uint4 *elems, uint64_t n_elems, uint64_t batch, uint4 *tree,
                         hipStream_t s) {
    if (leaf_m != 4 && leaf_m != 6) return B64_ERR_INVALID_ARGS;
    if (batch == 0 || n_elems % batch != 0 || batch > 0xFFFFFFFFull) return B64_ERR_INCORRECT_BATCH_SIZE;
    const uint64_t n_leaves = n_elems / batch;
    if (n_leaves == 0 || (n_leaves & (n_leaves - 1))) return B64_ERR_POWER_OF_TWO_LENGTH_REQUIRED;
    B64_TRY(vision::check_device());
    const uint4 *leaf = vision::device_data(leaf_m);
    const uint4 *node = vision::device_data(4);
    if (!leaf || !node) return B64_ERR_HIP;
    HIP_TRY(vision::launch_digest(leaf_m, leaf, elems, n_leaves, (uint32_t)batch, tree, s));
    HIP_TRY(vision::launch_tree_levels(node, tree, n_leaves, s));
    return B64_OK;
}

}  // namespace b64

extern "C" {

int b64_vision_permute(uint32_t m, void *d_states, uint64_t n, void *stream) {
    if ((m != 4 && m != 6) || (!d_states && n)) return B64_ERR_INVALID_ARGS;
    B64_TRY(b64::vision::check_device());
    const uint4 *dev = b64::vision::device_data((int)m);
    if (!dev) return B64_ERR_HIP;
    HIP_TRY(b64::vision::launch_permute((int)m, dev, (uint4 *)d_states, n, as_stream(stream)));
    return B64_OK;
}

int b64_vision_digest(uint32_t m, const void *d_msgs, uint64_t n, uint64_t msg_elems, void *d_out, void *stream) {
    if ((m != 4 && m != 6) || ((!d_msgs && n * msg_elems) || (!d_out && n)) || msg_elems > 0xFFFFFFFFull)
        return B64_ERR_INVALID_ARGS;
    B64_TRY(b64::vision::check_device());
    const uint4 *dev = b64::vision::device_data((int)m);
    if (!dev) return B64_ERR_HIP;
    HIP_TRY(b64::vision::launch_digest((int)m, dev, (const uint4 *)d_msgs, n, (uint32_t)msg_elems, (uint4 *)d_out,
                                       as_stream(stream)));
    return B64_OK;
}

int b64_merkle_commit_vision(uint32_t leaf_m, const void *d_elems, uint64_t n_elems, uint64_t batch, void *d_tree,
                             uint8_t root_out[32], void *stream) {
    if (!d_elems || !d_tree) return B64_ERR_INVALID_ARGS;
    const hipStream_t s = as_stream(stream);
    B64_TRY(b64::vision_merkle_commit((int)leaf_m, (const uint4 *)d_elems, n_elems, batch, (uint4 *)d_tree, s));
    const uint64_t n_leaves = n_elems / batch;
    if (root_out) {
        HIP_TRY(hipMemcpyAsync(root_out, (const uint8_t *)d_tree + 32 * (2 * n_leaves - 2), 32,
                               hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    return B64_OK;
}

}  // extern "C"
