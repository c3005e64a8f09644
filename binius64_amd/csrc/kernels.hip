// gfx950 kernels for the B128 encode / commit / fold path and their launchers.
#include <hip/hip_runtime.h>

#include <atomic>
#include <stdint.h>

#include "gf128_dev.hpp"
#include "internal.hpp"
#include "sha256_dev.hpp"

namespace b64 {

constexpr int BLOCK = 256;

static inline unsigned grid_for(uint64_t n, int per_block = BLOCK) {
    return (unsigned)((n + per_block - 1) / per_block);
}

// ---------------------------------------------------------------------------
// Element-wise products (device multiply parity surface)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_gf_mul(const uint4 *__restrict__ a, const uint4 *__restrict__ b,
                                                  uint4 *__restrict__ out, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) out[i] = gf_mul(a[i], b[i]);
}

__global__ __launch_bounds__(BLOCK) void k_gf_mul_const(const uint4 *__restrict__ a, uint4 c,
                                                        uint4 *__restrict__ out, uint64_t n) {
    __shared__ uint4 tab[WIN_ENTRIES];
    if (threadIdx.x < WIN_ROWS) win_build_row(tab, c, threadIdx.x);
    __syncthreads();
    for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * BLOCK)
        out[i] = win_mul(tab, a[i]);
}

hipError_t launch_gf_mul(const uint4 *a, const uint4 *b, uint4 *out, uint64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_gf_mul<<<grid_for(n), BLOCK, 0, s>>>(a, b, out, n);
    return hipGetLastError();
}

hipError_t launch_gf_mul_const(const uint4 *a, uint4 c, uint4 *out, uint64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    unsigned g = grid_for(n);
    if (g > 4096) g = 4096;
    k_gf_mul_const<<<g, BLOCK, 0, s>>>(a, c, out, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// encode_batch front end (reed_solomon.rs:152-187): out[r*2^m + i] = msg[rev_m(i)]
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_bitrev_replicate(const uint4 *__restrict__ msg,
                                                            uint4 *__restrict__ out, int m, int lir) {
    const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    const uint64_t n = 1ull << m;
    if (i >= n) return;
    const uint64_t src = m == 0 ? 0 : (__brevll(i) >> (64 - m));
    const uint4 v = msg[src];
    for (int r = 0; r < (1 << lir); r++) out[((uint64_t)r << m) | i] = v;
}

hipError_t launch_bitrev_replicate(const uint4 *msg, uint4 *out, int m, int lir, hipStream_t s) {
    k_bitrev_replicate<<<grid_for(1ull << m), BLOCK, 0, s>>>(msg, out, m, lir);
    return hipGetLastError();
}

// Pre-expanded twiddles (the GenericPreExpanded domain, ntt/domain_context.rs:
// 117-151): entry 2^layer + block is the XOR of W[layer][i] over the set bits i
// of block.
__global__ __launch_bounds__(BLOCK) void k_expand_twiddles(const uint4 *__restrict__ W, uint4 *__restrict__ X,
                                                           uint32_t n) {
    const uint32_t idx = blockIdx.x * BLOCK + threadIdx.x;
    if (idx >= n) return;
    uint4 acc = make_uint4(0, 0, 0, 0);
    if (idx) {
        const int layer = 31 - __clz(idx);
        const uint32_t block = idx - (1u << layer);
        const uint4 *Wl = W + layer * W_STRIDE;
        for (int i = 0; i < layer; i++)
            if ((block >> i) & 1) acc = gxor(acc, Wl[i]);
    }
    X[idx] = acc;
}

hipError_t launch_expand_twiddles(const uint4 *W, uint4 *X, int x_layers, hipStream_t s) {
    if (x_layers < 0 || x_layers > X_MAX_LAYERS) return hipErrorInvalidValue;
    const uint32_t n = 1u << x_layers;
    k_expand_twiddles<<<grid_for(n), BLOCK, 0, s>>>(W, X, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Additive NTT (ntt/reference.rs:17-76): LDS-tiled passes of up to NTT_MAX_K layers.
//
// A pass over layers [layer0, layer0+k) flips index bits lb+k-1 .. lb with
// lb = log_d - layer0 - k.  A tile holds every combination of those k bits and
// of the c lowest index bits (contiguous in HBM, so loads coalesce); the bits in
// between and above come from the tile id.  Twiddles only depend on bits above
// the flipped bit, so a tile needs 2^j distinct twiddles in its j-th layer.
// Where the low bits are too few to fill a tile (the last passes, lb small),
// the tile also takes the e bits just above the flipped ones: 2^e independent
// butterfly groups, and 2^(j+e) twiddles in layer j.
// ---------------------------------------------------------------------------
constexpr int NTT_TILE_LOG = 11;
constexpr int NTT_MAX_K = 7;

// Runtime knobs (set through b64_config_set): the multiply used by the NTT
// passes (0 = general product, 1 = LDS window tables where they pay, 2 = as 1
// with forward 4-layer passes run in registers, k_ntt_pass_r16), the
// largest number of layers per pass, the tile size, the butterflies per
// twiddle that justify a table, and the tables resident in LDS at once.
NttTuning g_ntt_tuning = {1, NTT_MAX_K, NTT_TILE_LOG, 128, 2, 0};
NttTuning g_ntt_tuning_size[MAX_DIM + 1] = {};

const NttTuning &ntt_tuning_for(int log_d) {
    static bool init = false;
    if (!init) {
        for (auto &t : g_ntt_tuning_size) t.variant = -1;
        init = true;
    }
    if (log_d >= 0 && log_d <= MAX_DIM && g_ntt_tuning_size[log_d].variant >= 0) return g_ntt_tuning_size[log_d];
    return g_ntt_tuning;
}

NttTuning g_batch_tuning = {1, 4, NTT_TILE_LOG, 128, 2, 0};
NttTuning g_batch_tuning_size[MAX_DIM + 1] = {};

const NttTuning &batch_tuning_for(int log_d) {
    static bool init = false;
    if (!init) {
        for (auto &t : g_batch_tuning_size) t.variant = -1;
        init = true;
    }
    if (log_d >= 0 && log_d <= MAX_DIM && g_batch_tuning_size[log_d].variant >= 0) return g_batch_tuning_size[log_d];
    return g_batch_tuning;
}


__device__ __forceinline__ uint32_t rev_bits(uint32_t x, int b) { return b ? (__brev(x) >> (32 - b)) : 0u; }

constexpr int NTT_PER_THREAD = (1 << NTT_TILE_LOG) / BLOCK;

// Block barrier that orders LDS only.  __syncthreads() also waits for every
// outstanding global load, which would stall on the next tile's prefetch; the
// NTT blocks share data through LDS alone.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Position of a tile: it covers index bits lb .. lb+k+e-1 (k flipped, e above
// them) and 0 .. c-1; the bits in between (mid) and above (hi) come from the
// tile id.  Tile element L = x << (k+c) | bf << c | low sits at index
// base | (L >> c) << lb | low.
struct NttTile {
    uint64_t base, mid, hi;
    int lb, n_el;
};
__device__ __forceinline__ NttTile ntt_tile_at(const NttPassArgs &a, uint64_t tile) {
    NttTile t;
    t.lb = a.log_d - a.layer0 - a.k;
    const int mid_bits = t.lb - a.c;
    t.mid = tile & ((1ull << mid_bits) - 1);
    t.hi = tile >> mid_bits;
    t.base = (t.hi << (t.lb + a.k + a.e)) | (t.mid << a.c);
    t.n_el = 1 << (a.k + a.c + a.e);
    return t;
}

// A tile in flight: thread x holds elements x, x + BLOCK, ... of the load order,
// as one register vector (the compiler keeps it in VGPRs where a uint4 array
// passed between functions can end up in scratch memory).
template <int NPT>
struct NttTileRegsN {
    typedef uint32_t vec __attribute__((ext_vector_type(4 * NPT)));
    static constexpr int n = NPT;
    vec v;
    __device__ __forceinline__ void set(int i, uint4 e) {
        v[4 * i] = e.x;
        v[4 * i + 1] = e.y;
        v[4 * i + 2] = e.z;
        v[4 * i + 3] = e.w;
    }
    __device__ __forceinline__ uint4 get(int i) const {
        return make_uint4(v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]);
    }
};
typedef NttTileRegsN<NTT_PER_THREAD> NttTileRegs;

// Issues the tile's loads.  Element L = (bf << c) | low of the tile holds
// index base | bf << lb | low.  In message mode the loads run in message order
// so they are runs of 2^k consecutive message elements: with
// i = bf << (m-k) | mid << c | low, rev_m(i) = rev_k(bf) | rev(mid) << k | rev_c(low) << (m-c).
template <int NPT>
__device__ __forceinline__ void ntt_fetch_tile(NttTileRegsN<NPT> &R, const NttPassArgs &a, const NttTile &t) {
    const int k = a.k, c = a.c;
    if (a.msg == nullptr) {
        const int cmask = (1 << c) - 1;
#pragma unroll
        for (int i = 0; i < NPT; i++) {
            // the test is uniform across the block; in tiles smaller than a
            // block, lanes past the end re-read a valid element (not stored)
            const int L = (threadIdx.x + i * BLOCK) & (t.n_el - 1);
            if (i * BLOCK < t.n_el) R.set(i, a.data[t.base | ((uint64_t)(L >> c) << t.lb) | (uint64_t)(L & cmask)]);
        }
    } else {
        const int mid_bits = t.lb - c;
        const uint64_t mbase = mid_bits ? ((__brevll(t.mid) >> (64 - mid_bits)) << k) : 0;
        const int kmask = (1 << k) - 1;
#pragma unroll
        for (int i = 0; i < NPT; i++) {
            const int q = (threadIdx.x + i * BLOCK) & (t.n_el - 1);
            if (i * BLOCK < t.n_el) R.set(i, a.msg[mbase + (q & kmask) + ((uint64_t)(q >> k) << (a.m - c))]);
        }
    }
}

// Writes a fetched tile into LDS in tile order.
template <int NPT>
__device__ __forceinline__ void ntt_put_tile(uint4 *__restrict__ s, const NttTileRegsN<NPT> &R, const NttPassArgs &a,
                                             int n_el) {
    const int k = a.k, c = a.c;
    if (a.msg == nullptr) {
#pragma unroll
        for (int i = 0; i < NPT; i++) {
            const int L = threadIdx.x + i * BLOCK;
            if (L < n_el) s[L] = R.get(i);
        }
    } else {
        const int kmask = (1 << k) - 1;
#pragma unroll
        for (int i = 0; i < NPT; i++) {
            const int q = threadIdx.x + i * BLOCK;
            if (q < n_el) s[(rev_bits(q & kmask, k) << c) | rev_bits(q >> k, c)] = R.get(i);
        }
    }
}

// Every thread issues all of its loads before the first LDS store, so a block
// keeps 2^(k+c) loads in flight.
__device__ __forceinline__ void ntt_load_tile(uint4 *__restrict__ s, const NttPassArgs &a, const NttTile &t) {
    NttTileRegs R;
    ntt_fetch_tile(R, a, t);
    ntt_put_tile(s, R, a, t.n_el);
}

template <int NPT = NTT_PER_THREAD>
__device__ __forceinline__ void ntt_store_tile(const uint4 *__restrict__ s, const NttPassArgs &a, const NttTile &t) {
    const int cmask = (1 << a.c) - 1;
    NttTileRegsN<NPT> R;
#pragma unroll
    for (int i = 0; i < NPT; i++)
        if (i * BLOCK < t.n_el) R.set(i, s[(threadIdx.x + i * BLOCK) & (t.n_el - 1)]);
#pragma unroll
    for (int i = 0; i < NPT; i++) {
        const int L = threadIdx.x + i * BLOCK;
        if (L < t.n_el) a.data[t.base | ((uint64_t)(L >> a.c) << t.lb) | (uint64_t)(L & cmask)] = R.get(i);
    }
}

// Twiddles of every layer of the pass for this tile:
// tw[(2^j - 1) << e | B] = twiddle(layer0 + j, hi << (j+e) | B) = XOR of
// W[layer][i] over the set bits i of the block index, for the 2^(j+e) blocks B
// of layer j in the tile.  The part from hi (bits above the tile) is shared by
// a layer's twiddles and is summed once per layer, with the loads of a layer
// issued together.  Needs k + e <= NTT_MAX_K.  Ends with a barrier.
constexpr int NTT_TW = (1 << NTT_MAX_K) - 1;
__host__ __device__ __forceinline__ int ntt_tw_base(int j, int e) { return ((1 << j) - 1) << e; }

__device__ __forceinline__ void ntt_tile_twiddles(const NttPassArgs &a, uint64_t hi, uint4 *__restrict__ tw,
                                                  uint4 *__restrict__ thi) {
    const int k = a.k, e = a.e;
    if ((int)threadIdx.x < k) {
        const int j = threadIdx.x;
        const uint4 *Wl = a.W + (a.layer0 + j) * W_STRIDE + j + e;
        const int n_hi = a.layer0 - e;
        uint4 acc = make_uint4(0, 0, 0, 0);
        for (int i0 = 0; i0 < n_hi; i0 += 8) {
            uint4 w[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int i = i0 + u;
                w[u] = (i < n_hi && ((hi >> i) & 1)) ? Wl[i] : make_uint4(0, 0, 0, 0);
            }
#pragma unroll
            for (int u = 0; u < 8; u++) acc = gxor(acc, w[u]);
        }
        thi[j] = acc;
    }
    lds_barrier();
    for (int t = threadIdx.x; t < ntt_tw_base(k, e); t += BLOCK) {
        const int j = 31 - __clz((t >> e) + 1);  // (t >> e) + 1 in [2^j, 2^(j+1))
        const int loc = t - ntt_tw_base(j, e);    // block index, j + e bits
        const uint4 *Wl = a.W + (a.layer0 + j) * W_STRIDE;
        uint4 w[NTT_MAX_K - 1];
#pragma unroll
        for (int i = 0; i < NTT_MAX_K - 1; i++)
            w[i] = (i < j + e && ((loc >> i) & 1)) ? Wl[i] : make_uint4(0, 0, 0, 0);
        uint4 acc = thi[j];
#pragma unroll
        for (int i = 0; i < NTT_MAX_K - 1; i++) acc = gxor(acc, w[i]);
        tw[t] = acc;
    }
    lds_barrier();
}

template <bool INVERSE>
__global__ __launch_bounds__(BLOCK) void k_ntt_pass(NttPassArgs a) {
    __shared__ uint4 s[1 << NTT_TILE_LOG];
    __shared__ uint4 tw[NTT_TW];
    __shared__ uint4 thi[NTT_MAX_K];
    const int k = a.k, c = a.c;
    const NttTile T = ntt_tile_at(a, blockIdx.x);
    const int n_el = T.n_el;

    ntt_load_tile(s, a, T);
    ntt_tile_twiddles(a, T.hi, tw, thi);

    for (int jj = 0; jj < k; jj++) {
        const int j = INVERSE ? (k - 1 - jj) : jj;
        const uint4 *twl = tw + ntt_tw_base(j, a.e);
        const int pos = c + (k - 1 - j);
        const int lowmask = (1 << pos) - 1;
        for (int P = threadIdx.x; P < (n_el >> 1); P += BLOCK) {
            const int L0 = ((P >> pos) << (pos + 1)) | (P & lowmask);
            const int L1 = L0 | (1 << pos);
            const uint4 t = twl[P >> pos];
            uint4 u = s[L0], v = s[L1];
            if (!INVERSE) {
                u = gxor(u, gf_mul(v, t));
                v = gxor(v, u);
            } else {
                v = gxor(v, u);
                u = gxor(u, gf_mul(v, t));
            }
            s[L0] = u;
            s[L1] = v;
        }
        lds_barrier();
    }

    ntt_store_tile(s, a, T);
}

// Same pass structure, but each butterfly multiplies by its block's twiddle
// through an LDS window table.  The 2^j blocks of layer j are contiguous ranges
// of the butterfly enumeration; up to G tables are built at a time (one lane
// per table row, all lanes of the block) and consumed by the butterflies of
// those blocks.  Layers with fewer than tab_min butterflies per twiddle use the
// general product instead.
template <int G, int TL = NTT_TILE_LOG>
struct alignas(256) NttTabSmem {
    uint4 s[1 << TL];
    uint4 tabs[G * WIN_ENTRIES];
    uint4 tw[NTT_TW];
    uint4 thi[NTT_MAX_K];
};

// The k layers of a pass on a tile resident in sm.s, twiddles in sm.tw.
// HYBRID = false compiles the table path alone (the launcher picks it when every
// layer of the pass qualifies), which keeps the general multiply's registers out
// of the kernel.
template <bool INVERSE, int G, bool HYBRID, class SM>
__device__ __forceinline__ void ntt_tab_layers(const NttPassArgs &a, int n_el, SM &sm) {
    uint4 *__restrict__ s = sm.s;
    uint4 *__restrict__ tabs = sm.tabs;
    const int k = a.k, c = a.c;
    for (int jj = 0; jj < k; jj++) {
        const int j = INVERSE ? (k - 1 - jj) : jj;
        const uint4 *__restrict__ tw = sm.tw + ntt_tw_base(j, a.e);
        const int pos = c + (k - 1 - j);
        const int per_blk = 1 << pos;
        const int lowmask = per_blk - 1;
        const int nblk = 1 << (j + a.e);
        if (HYBRID && per_blk < a.tab_min) {
            // few butterflies per twiddle: a table would not pay for itself
            for (int P = threadIdx.x; P < (n_el >> 1); P += BLOCK) {
                const int L0 = ((P >> pos) << (pos + 1)) | (P & lowmask);
                const int L1 = L0 | per_blk;
                const uint4 t = tw[P >> pos];
                uint4 u = s[L0], v = s[L1];
                if (!INVERSE) {
                    u = gxor(u, gf_mul(v, t));
                    v = gxor(v, u);
                } else {
                    v = gxor(v, u);
                    u = gxor(u, gf_mul(v, t));
                }
                s[L0] = u;
                s[L1] = v;
            }
            lds_barrier();
            continue;
        }
        for (int g0 = 0; g0 < nblk; g0 += G) {
            const int ng = nblk - g0 < G ? nblk - g0 : G;
            for (int r = threadIdx.x; r < WIN_ROWS * ng; r += BLOCK)
                win_build_row(tabs + (r >> 5) * WIN_ENTRIES, tw[g0 + (r >> 5)], r & 31);
            lds_barrier();
            const int n_bf = ng * per_blk;
            for (int q = threadIdx.x; q < n_bf; q += BLOCK) {
                const int P = (g0 << pos) + q;
                const int L0 = ((P >> pos) << (pos + 1)) | (P & lowmask);
                const int L1 = L0 | per_blk;
                const uint4 *tab = tabs + (q >> pos) * WIN_ENTRIES;
                uint4 u = s[L0], v = s[L1];
                if (!INVERSE) {
                    u = win_mul_perm(lds_addr(tab), v, u);
                    v = gxor(v, u);
                } else {
                    v = gxor(v, u);
                    u = win_mul_perm(lds_addr(tab), v, u);
                }
                s[L0] = u;
                s[L1] = v;
            }
            lds_barrier();
        }
    }

}

template <bool INVERSE, int G, bool HYBRID, class SM>
__device__ __forceinline__ void ntt_tab_tile(const NttPassArgs &a, uint64_t tile, SM &sm) {
    const NttTile T = ntt_tile_at(a, tile);
    ntt_load_tile(sm.s, a, T);
    ntt_tile_twiddles(a, T.hi, sm.tw, sm.thi);
    ntt_tab_layers<INVERSE, G, HYBRID>(a, T.n_el, sm);
    ntt_store_tile(sm.s, a, T);
}

template <bool INVERSE, int G, bool HYBRID>
__global__ __launch_bounds__(BLOCK) void k_ntt_pass_tab(NttPassArgs a) {
    __shared__ NttTabSmem<G> sm;
    ntt_tab_tile<INVERSE, G, HYBRID>(a, blockIdx.x, sm);
}

// Tables-only passes are latency bound on LDS lookups.  This form takes tiles of
// at most 2^10 elements, so its LDS fits four blocks per CU, and asks for four
// waves per SIMD (at most 128 VGPRs).
constexpr int NTT_W4_TILE_LOG = 10;
template <bool INVERSE, int G>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void k_ntt_pass_tab_w4(NttPassArgs a) {
    __shared__ NttTabSmem<G, NTT_W4_TILE_LOG> sm;
    ntt_tab_tile<INVERSE, G, false>(a, blockIdx.x, sm);
}

template <bool INVERSE, bool HYBRID>
static void launch_ntt_tab_h(const NttPassArgs &a, uint32_t n_tiles, int tab_group, hipStream_t s) {
    switch (tab_group) {
        case 8: k_ntt_pass_tab<INVERSE, 8, HYBRID><<<n_tiles, BLOCK, 0, s>>>(a); break;
        case 4: k_ntt_pass_tab<INVERSE, 4, HYBRID><<<n_tiles, BLOCK, 0, s>>>(a); break;
        case 1: k_ntt_pass_tab<INVERSE, 1, HYBRID><<<n_tiles, BLOCK, 0, s>>>(a); break;
        default: k_ntt_pass_tab<INVERSE, 2, HYBRID><<<n_tiles, BLOCK, 0, s>>>(a); break;
    }
}

template <bool INVERSE>
static void launch_ntt_tab_w4(const NttPassArgs &a, uint32_t n_tiles, int tab_group, hipStream_t s) {
    switch (tab_group) {
        case 8: k_ntt_pass_tab_w4<INVERSE, 8><<<n_tiles, BLOCK, 0, s>>>(a); break;
        case 4: k_ntt_pass_tab_w4<INVERSE, 4><<<n_tiles, BLOCK, 0, s>>>(a); break;
        case 1: k_ntt_pass_tab_w4<INVERSE, 1><<<n_tiles, BLOCK, 0, s>>>(a); break;
        default: k_ntt_pass_tab_w4<INVERSE, 2><<<n_tiles, BLOCK, 0, s>>>(a); break;
    }
}

// Four-layer passes in registers (variant 2).  A thread owns the 16 elements
// that differ only in the pass's flipped bits lb .. lb+3 and runs all four
// layers on them in registers; the 15 twiddles of the pass depend only on the
// bits above lb+3 (hi), so a block that stays inside one hi group builds the
// 15 window tables (120 KiB of LDS) once and streams 1024-position slices
// through them: no LDS traffic for the data, no block barrier between layers.
// Used where a hi group has at least R16_BLOCK positions (lb >= 10), forward
// only; the message-gathering first pass of an encode reads, per thread, one
// 16-element run of the message (the flipped bits are the low 4 bits of the
// bit-reversed message index).
constexpr int R16_BLOCK = 1024;
constexpr int R16_MIN_LB = 10;
constexpr int R16_MAX_LB = 32;  // lane offsets within a hi group are 32-bit
constexpr int R16_TABLES = 15;
constexpr int R16_LOOKUPS = 4;  // table reads in flight per wave (VGPR budget: 128)
struct alignas(256) R16Smem {
    uint4 tabs[R16_TABLES * WIN_ENTRIES];
    uint4 tw[NTT_TW];
    uint4 thi[NTT_MAX_K];
};

// Element e (0..15, the value of index bits lb .. lb+3) of lane offset `low` in
// hi group `hi`.  For the message-gathering pass, codeword index bits
// lb .. lb+3 are message bits 3 .. 0 (reversed), so a lane's 16 elements are
// one 16-element run of the message.
template <bool MSG>
__device__ __forceinline__ uint4 r16_load(const NttPassArgs &a, uint64_t hi, int lb, uint64_t low, int e) {
    if (MSG) {
        const int mb = a.m - 4;
        const uint64_t run = mb ? (__brevll(low) >> (64 - mb)) << 4 : 0;
        return a.msg[run + rev_bits(e, 4)];
    }
    return (a.data + (hi << (lb + 4)) + ((uint64_t)e << lb))[(uint32_t)low];
}

// Block `blk` of the pass (single-job launches: blockIdx.x; multi-job: the
// block's index within its job).
template <bool MSG>
__device__ __forceinline__ void ntt_r16_core(const NttPassArgs &a, uint64_t per_block, uint64_t blk, R16Smem &sm) {
    const int lb = a.log_d - a.layer0 - 4;
    const uint64_t p0 = blk * per_block;
    const uint64_t hi = p0 >> lb;
    ntt_tile_twiddles(a, hi, sm.tw, sm.thi);  // a.k = 4, a.e = 0; ends with a barrier
    for (int r = threadIdx.x; r < R16_TABLES * WIN_ROWS; r += R16_BLOCK)
        win_build_row(sm.tabs + (r >> 5) * WIN_ENTRIES, sm.tw[r >> 5], r & 31);
    lds_barrier();
    const uint32_t tabs = lds_addr(sm.tabs);
    const uint64_t lowmask = (1ull << lb) - 1;
    // Every thread makes per_block / R16_BLOCK passes of the loop (per_block is a
    // multiple of R16_BLOCK), so `more` is uniform across the block.  The loads of
    // the next position go out during the last layer, pair by pair as the
    // current position's elements are stored, so each wave's HBM traffic
    // overlaps its own arithmetic instead of running as a separate phase.
    // lane offsets within the hi group: low0 + i for i = threadIdx.x, +R16_BLOCK, ...
    const uint32_t low0 = (uint32_t)(p0 & lowmask), nper = (uint32_t)per_block;
    uint32_t i = threadIdx.x;
    uint4 x[16];
#pragma unroll
    for (int e = 0; e < 16; e++) x[e] = r16_load<MSG>(a, hi, lb, low0 + i, e);
    for (;;) {
        const uint32_t in = i + R16_BLOCK;
        const bool more = in < nper;
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int bit = 3 - j;  // layer j flips element bit 3 - j
#pragma unroll
            for (int e = 0; e < 16; e++) {
                if (e & (1 << bit)) continue;
                const int blk = e >> (bit + 1);
                const uint32_t tb = tabs + (uint32_t)(((1 << j) - 1 + blk) * WIN_ENTRIES * 16);
                uint4 u = x[e], v = x[e | (1 << bit)];
                u = gxor(u, win_mul_perm<R16_LOOKUPS>(tb, v));
                v = gxor(v, u);
                x[e] = u;
                x[e | (1 << bit)] = v;
            }
        }
        // layer 3 (element bit 0): each finished pair is stored, and its
        // registers take the next position's pair (after the last position, a
        // re-read of the pair just stored, unused: no branch in the loop body)
        const uint32_t low = low0 + i, lown = more ? low0 + in : low;
#pragma unroll
        for (int e = 0; e < 16; e += 2) {
            const uint32_t tb = tabs + (uint32_t)((7 + (e >> 1)) * WIN_ENTRIES * 16);
            uint4 u = x[e], v = x[e + 1];
            u = gxor(u, win_mul_perm<R16_LOOKUPS>(tb, v));
            v = gxor(v, u);
            uint4 *row = a.data + (hi << (lb + 4));
            row[((uint64_t)e << lb) + low] = u;
            row[((uint64_t)(e + 1) << lb) + low] = v;
            x[e] = r16_load<MSG>(a, hi, lb, lown, e);
            x[e + 1] = r16_load<MSG>(a, hi, lb, lown, e + 1);
        }
        if (!more) break;
        i = in;
    }
}

template <bool MSG>
__global__ __launch_bounds__(R16_BLOCK) void k_ntt_pass_r16(NttPassArgs a, uint64_t per_block) {
    __shared__ R16Smem sm;
    ntt_r16_core<MSG>(a, per_block, blockIdx.x, sm);
}

static uint32_t device_cus() {
    static std::atomic<int> cached[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    int cus = cached[dev].load(std::memory_order_relaxed);
    if (cus == 0) {
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 1;
        cached[dev].store(cus, std::memory_order_relaxed);
    }
    return (uint32_t)cus;
}

// One block per CU (the tables take 120 KiB); positions per block a power of
// two that divides 2^lb, so a block never leaves its hi group.
static int r16_per_log(int log_d, int layer0, uint32_t cus) {
    const int lb = log_d - layer0 - 4;
    const int total_log = log_d - 4;
    int cus_log = 0;
    while ((1u << cus_log) < cus) cus_log++;
    int per_log = total_log - cus_log;
    if (per_log < R16_MIN_LB) per_log = R16_MIN_LB;
    if (per_log > lb) per_log = lb;
    return per_log;
}

static void launch_ntt_r16(const NttPassArgs &a, hipStream_t s) {
    const int per_log = r16_per_log(a.log_d, a.layer0, device_cus());
    const uint32_t blocks = 1u << (a.log_d - 4 - per_log);
    if (a.msg)
        k_ntt_pass_r16<true><<<blocks, R16_BLOCK, 0, s>>>(a, 1ull << per_log);
    else
        k_ntt_pass_r16<false><<<blocks, R16_BLOCK, 0, s>>>(a, 1ull << per_log);
}

// The deepest layer of a pass has the fewest butterflies per twiddle, 2^c.
// ---- passes through 2 KiB tables (variants 3 and 4) ---------------------------
// LDS-tile pass for the last layers of a transform, where a twiddle covers only
// 16..128 butterflies: every twiddle of a layer in the tile gets a 2 KiB table
// (w8_build_half, all threads of the block building at once, up to T8_GROUP
// tables per round), then every butterfly of the layer runs through them.
// Layers with fewer than tab_min butterflies per twiddle use the general product.
constexpr int T8_GROUP = 16;  // tables per build round (32 KiB)
template <int TL>
struct alignas(256) NttT8Smem {
    uint4 tabs[T8_GROUP * W8_ENTRIES];
    uint4 s[1 << TL];
    uint4 tw[NTT_TW];
    uint4 thi[NTT_MAX_K];
};

// Variant with every table of the tile built in one round: 1 KiB tables (rows
// 0..3, w4_mul) at the twiddle's index in tw, for tiles with at most T4U_TABLES
// twiddles (4 layers with e <= 1).
constexpr int T4U_TABLES = 30;
template <int TL>
struct alignas(256) NttT4USmem {
    uint4 tabs[T4U_TABLES * W4_ENTRIES];
    uint4 s[1 << TL];
    uint4 tw[T4U_TABLES];
    uint4 thi[NTT_MAX_K];
};

// With pre-expanded twiddles the tile's twiddle loads go out first, its element
// loads behind them, and the tables are built (a row per lane) while the
// element loads are in flight.
template <int TL>
__device__ __forceinline__ void ntt_t4u_core(const NttPassArgs &a, uint64_t tile, NttT4USmem<TL> &sm) {
    constexpr int NPT = (1 << TL) / BLOCK;
    const NttTile T = ntt_tile_at(a, tile);
    const int k = a.k, c = a.c, e = a.e, n_el = T.n_el;
    const int n_tab = ntt_tw_base(k, e);
    const bool pre = a.X != nullptr && a.layer0 + k <= a.x_layers;  // uniform
    uint4 twv = make_uint4(0, 0, 0, 0);
    if (pre && (int)threadIdx.x < n_tab) {
        const int t = threadIdx.x;
        const int j = 31 - __clz((t >> e) + 1);  // layer of tw[t]
        const uint64_t blk = (T.hi << (j + e)) | (uint64_t)(t - ntt_tw_base(j, e));
        twv = a.X[(1ull << (a.layer0 + j)) + blk];
    }
    NttTileRegsN<NPT> R;
    ntt_fetch_tile(R, a, T);
    if (pre) {
        if ((int)threadIdx.x < n_tab) sm.tw[threadIdx.x] = twv;
        lds_barrier();
    } else {
        ntt_tile_twiddles(a, T.hi, sm.tw, sm.thi);  // ends with a barrier
    }
    for (int r = threadIdx.x; r < W4_ROWS * n_tab; r += BLOCK)
        w8_build_row(sm.tabs + (r / W4_ROWS) * W4_ENTRIES, sm.tw[r / W4_ROWS], r % W4_ROWS);
    ntt_put_tile(sm.s, R, a, n_el);
    lds_barrier();
    uint4 *__restrict__ s = sm.s;
    const uint32_t tabs = lds_addr(sm.tabs);
    for (int j = 0; j < k; j++) {
        const int tw0 = ntt_tw_base(j, a.e);
        const int pos = c + (k - 1 - j);
        const int per_blk = 1 << pos;
        const int lowmask = per_blk - 1;
        for (int P = threadIdx.x; P < (n_el >> 1); P += BLOCK) {
            const int L0 = ((P >> pos) << (pos + 1)) | (P & lowmask);
            const int L1 = L0 | per_blk;
            uint4 u = s[L0], v = s[L1];
            u = w4_mul(tabs + (uint32_t)((tw0 + (P >> pos)) * W4_ENTRIES * 16), v, u);
            v = gxor(v, u);
            s[L0] = u;
            s[L1] = v;
        }
        lds_barrier();
    }
    ntt_store_tile<NPT>(sm.s, a, T);
}

template <int TL>
__global__ __launch_bounds__(BLOCK) void k_ntt_pass_t4u(NttPassArgs a) {
    __shared__ NttT4USmem<TL> sm;
    ntt_t4u_core<TL>(a, blockIdx.x, sm);
}

template <int TL>
__global__ __launch_bounds__(BLOCK) void k_ntt_pass_t8(NttPassArgs a) {
    __shared__ NttT8Smem<TL> sm;
    const NttTile T = ntt_tile_at(a, blockIdx.x);
    ntt_load_tile(sm.s, a, T);
    ntt_tile_twiddles(a, T.hi, sm.tw, sm.thi);  // ends with a barrier
    uint4 *__restrict__ s = sm.s;
    const uint32_t tabs = lds_addr(sm.tabs);
    const int k = a.k, c = a.c, n_el = T.n_el;
    for (int j = 0; j < k; j++) {
        const uint4 *__restrict__ tw = sm.tw + ntt_tw_base(j, a.e);
        const int pos = c + (k - 1 - j);
        const int per_blk = 1 << pos;
        const int lowmask = per_blk - 1;
        const int nblk = 1 << (j + a.e);
        if (per_blk < a.tab_min) {
            for (int P = threadIdx.x; P < (n_el >> 1); P += BLOCK) {
                const int L0 = ((P >> pos) << (pos + 1)) | (P & lowmask);
                const int L1 = L0 | per_blk;
                uint4 u = s[L0], v = s[L1];
                u = gxor(u, gf_mul(v, tw[P >> pos]));
                v = gxor(v, u);
                s[L0] = u;
                s[L1] = v;
            }
            lds_barrier();
            continue;
        }
        for (int g0 = 0; g0 < nblk; g0 += T8_GROUP) {
            const int ng = nblk - g0 < T8_GROUP ? nblk - g0 : T8_GROUP;
            for (int r = threadIdx.x; r < 16 * ng; r += BLOCK)
                w8_build_half(sm.tabs + (r >> 4) * W8_ENTRIES, tw[g0 + (r >> 4)], (r >> 1) & 7, r & 1);
            lds_barrier();
            const int n_bf = ng * per_blk;
            for (int q = threadIdx.x; q < n_bf; q += BLOCK) {
                const int P = (g0 << pos) + q;
                const int L0 = ((P >> pos) << (pos + 1)) | (P & lowmask);
                const int L1 = L0 | per_blk;
                uint4 u = s[L0], v = s[L1];
                u = w8_mul(tabs + (uint32_t)((q >> pos) * W8_ENTRIES * 16), v, u);
                v = gxor(v, u);
                s[L0] = u;
                s[L1] = v;
            }
            lds_barrier();
        }
    }
    ntt_store_tile(sm.s, a, T);
}

// Four-layer register pass through 2 KiB tables: as k_ntt_pass_r16, with 256-thread
// blocks and the 15 tables in 30 KiB, so hi groups of 256 positions (lb >= 8)
// qualify and several blocks share a CU.
constexpr int R8_BLOCK = 256;
constexpr int R8_MIN_LB = 8;
constexpr int R8_LOOKUPS = 4;  // table reads in flight per wave (VGPR budget: 128)
struct alignas(256) R8Smem {
    uint4 tabs[R16_TABLES * W8_ENTRIES];
    uint4 tw[NTT_TW];
    uint4 thi[NTT_MAX_K];
};

template <bool MSG>
__device__ __forceinline__ void ntt_r16w8_core(const NttPassArgs &a, uint64_t per_block, uint64_t blk, R8Smem &sm) {
    const int lb = a.log_d - a.layer0 - 4;
    const uint64_t p0 = blk * per_block;
    const uint64_t hi = p0 >> lb;
    ntt_tile_twiddles(a, hi, sm.tw, sm.thi);  // a.k = 4, a.e = 0; ends with a barrier
    for (int r = threadIdx.x; r < R16_TABLES * 16; r += R8_BLOCK)
        w8_build_half(sm.tabs + (r >> 4) * W8_ENTRIES, sm.tw[r >> 4], (r >> 1) & 7, r & 1);
    lds_barrier();
    const uint32_t tabs = lds_addr(sm.tabs);
    const uint64_t lowmask = (1ull << lb) - 1;
    // the loop of ntt_r16_core: the next position's loads go out during the last
    // layer (per_block is a multiple of R8_BLOCK, so `more` is uniform)
    const uint32_t low0 = (uint32_t)(p0 & lowmask), nper = (uint32_t)per_block;
    uint32_t i = threadIdx.x;
    uint4 x[16];
#pragma unroll
    for (int e = 0; e < 16; e++) x[e] = r16_load<MSG>(a, hi, lb, low0 + i, e);
    for (;;) {
        const uint32_t in = i + R8_BLOCK;
        const bool more = in < nper;
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int bit = 3 - j;
#pragma unroll
            for (int e = 0; e < 16; e++) {
                if (e & (1 << bit)) continue;
                const int blk = e >> (bit + 1);
                const uint32_t tb = tabs + (uint32_t)(((1 << j) - 1 + blk) * W8_ENTRIES * 16);
                uint4 u = x[e], v = x[e | (1 << bit)];
                u = gxor(u, w8_mul<R8_LOOKUPS>(tb, v));
                v = gxor(v, u);
                x[e] = u;
                x[e | (1 << bit)] = v;
            }
        }
        const uint32_t low = low0 + i, lown = more ? low0 + in : low;
#pragma unroll
        for (int e = 0; e < 16; e += 2) {
            const uint32_t tb = tabs + (uint32_t)((7 + (e >> 1)) * W8_ENTRIES * 16);
            uint4 u = x[e], v = x[e + 1];
            u = gxor(u, w8_mul<R8_LOOKUPS>(tb, v));
            v = gxor(v, u);
            uint4 *row = a.data + (hi << (lb + 4));
            row[((uint64_t)e << lb) + low] = u;
            row[((uint64_t)(e + 1) << lb) + low] = v;
            x[e] = r16_load<MSG>(a, hi, lb, lown, e);
            x[e + 1] = r16_load<MSG>(a, hi, lb, lown, e + 1);
        }
        if (!more) break;
        i = in;
    }
}

template <bool MSG>
__global__ __launch_bounds__(R8_BLOCK) void k_ntt_pass_r16w8(NttPassArgs a, uint64_t per_block) {
    __shared__ R8Smem sm;
    ntt_r16w8_core<MSG>(a, per_block, blockIdx.x, sm);
}

// Positions per r16w8 block: a power of two within the hi group, at least a
// block's worth, aiming at four blocks per CU.
static int r8_per_log(int log_d, int layer0, uint32_t cus) {
    const int lb = log_d - layer0 - 4;
    const int total_log = log_d - 4;
    int cus_log = 0;
    while ((1u << cus_log) < cus) cus_log++;
    int per_log = total_log - cus_log - 2;
    if (per_log < R8_MIN_LB) per_log = R8_MIN_LB;
    if (per_log > lb) per_log = lb;
    return per_log;
}

static void launch_ntt_r16w8(const NttPassArgs &a, hipStream_t s) {
    const int per_log = r8_per_log(a.log_d, a.layer0, device_cus());
    const uint32_t blocks = 1u << (a.log_d - 4 - per_log);
    if (a.msg)
        k_ntt_pass_r16w8<true><<<blocks, R8_BLOCK, 0, s>>>(a, 1ull << per_log);
    else
        k_ntt_pass_r16w8<false><<<blocks, R8_BLOCK, 0, s>>>(a, 1ull << per_log);
}

static void launch_ntt_t4u(const NttPassArgs &a, uint32_t n_tiles, hipStream_t s) {
    if (a.k + a.c + a.e <= 9)
        k_ntt_pass_t4u<9><<<n_tiles, BLOCK, 0, s>>>(a);
    else
        k_ntt_pass_t4u<10><<<n_tiles, BLOCK, 0, s>>>(a);
}

static void launch_ntt_t8(const NttPassArgs &a, uint32_t n_tiles, hipStream_t s) {
    const int tl = a.k + a.c + a.e;
    if (tl <= 9)
        k_ntt_pass_t8<9><<<n_tiles, BLOCK, 0, s>>>(a);
    else if (tl == 10)
        k_ntt_pass_t8<10><<<n_tiles, BLOCK, 0, s>>>(a);
    else
        k_ntt_pass_t8<11><<<n_tiles, BLOCK, 0, s>>>(a);
}

// Kernel of one pass.  The register pass addresses a hi group's elements with
// 32-bit lane offsets below 2^lb, so it takes lb <= R16_MAX_LB only.
int ntt_pass_kernel(const NttTuning &tu, int log_d, int layer0, int k, int c, int e, bool inverse) {
    const int lb = log_d - layer0 - k;
    const bool reg4 = !inverse && k == 4 && e == 0 && lb <= R16_MAX_LB;
    if (reg4 && (tu.variant == 2 || tu.variant == 3) && lb >= R16_MIN_LB) return B64_NTT_KERNEL_R16;
    if (reg4 && tu.variant >= 3 && lb >= R8_MIN_LB) return B64_NTT_KERNEL_R16W8;
    if (tu.variant >= 3 && !inverse && k + c + e <= 11) {
        if (tu.tab_group == 1 && ntt_tw_base(k, e) <= T4U_TABLES && k + c + e <= 10) return B64_NTT_KERNEL_T4U;
        return B64_NTT_KERNEL_T8;
    }
    if (tu.variant >= 1) {
        if ((1 << c) >= tu.tab_min)
            return (tu.waves == 4 && k + c + e <= NTT_W4_TILE_LOG) ? B64_NTT_KERNEL_TAB_W4 : B64_NTT_KERNEL_TAB;
        return B64_NTT_KERNEL_TAB_HYBRID;
    }
    return B64_NTT_KERNEL_GENERAL;
}

// Work items (blocks x threads) of that pass's grid; HIP launches take fewer
// than 2^32.
uint64_t ntt_pass_workitems(int kernel, int log_d, int layer0, int k, int c, int e, uint32_t cus) {
    if (kernel == B64_NTT_KERNEL_R16)
        return (1ull << (log_d - 4 - r16_per_log(log_d, layer0, cus))) * (uint64_t)R16_BLOCK;
    if (kernel == B64_NTT_KERNEL_R16W8)
        return (1ull << (log_d - 4 - r8_per_log(log_d, layer0, cus))) * (uint64_t)R8_BLOCK;
    return (1ull << (log_d - k - c - e)) * (uint64_t)BLOCK;
}

uint32_t ntt_device_cus() { return device_cus(); }

template <bool INVERSE>
static void launch_ntt_tab(const NttPassArgs &a, uint32_t n_tiles, int kernel, const NttTuning &tu, hipStream_t s) {
    if (kernel == B64_NTT_KERNEL_TAB_W4)
        launch_ntt_tab_w4<INVERSE>(a, n_tiles, tu.tab_group, s);
    else if (kernel == B64_NTT_KERNEL_TAB)
        launch_ntt_tab_h<INVERSE, false>(a, n_tiles, tu.tab_group, s);
    else
        launch_ntt_tab_h<INVERSE, true>(a, n_tiles, tu.tab_group, s);
}


constexpr int NTT_MIN_TILES_LOG = 10;

int ntt_pass_plan(const NttTuning &tu, int log_d, int skip_early, int skip_late, bool msg_first, int *layer0,
                  int *kk, int *cc, int *ee) {
    const int first = skip_early, end = log_d - skip_late;
    const int n_layers = end - first;
    if (n_layers <= 0) return 0;
    int max_k = tu.max_k;
    if (max_k < 1 || max_k > NTT_MAX_K) max_k = NTT_MAX_K;
    int tile_log = tu.tile_log;
    if (tile_log < max_k || tile_log > NTT_TILE_LOG) tile_log = NTT_TILE_LOG;
    const int n_passes = (n_layers + max_k - 1) / max_k;
    // balanced split of the layer range into passes
    int at = first;
    for (int p = 0; p < n_passes; p++) {
        const int left = end - at, passes_left = n_passes - p;
        const int k = (left + passes_left - 1) / passes_left;
        const int lb = log_d - at - k;
        layer0[p] = at;
        kk[p] = k;
        cc[p] = lb < (tile_log - k) ? lb : (tile_log - k);
        // fill the tile from the bits above the flipped ones (there are `at`),
        // as long as 2^NTT_MIN_TILES_LOG tiles remain to spread over the device
        int e = tile_log - k - cc[p];
        if (e > at) e = at;
        if (e > NTT_MAX_K - k) e = NTT_MAX_K - k;
        const int spare = log_d - k - cc[p] - NTT_MIN_TILES_LOG;
        if (e > spare) e = spare > 0 ? spare : 0;
        ee[p] = (p == 0 && msg_first) ? 0 : e;
        at += k;
    }
    return n_passes;
}

hipError_t launch_ntt(const DeviceDomain &dd, uint4 *data, int log_d, int skip_early, int skip_late,
                      bool inverse, hipStream_t s, const uint4 *msg, int msg_log) {
    // message mode needs the first pass to start at layer lir = log_d - msg_log
    if (msg && (inverse || skip_early != log_d - msg_log)) return hipErrorInvalidValue;
    int starts[64], ks[64], cs[64], es[64];
    const NttTuning &tu = ntt_tuning_for(log_d);
    const int n_passes = ntt_pass_plan(tu, log_d, skip_early, skip_late, msg != nullptr, starts, ks, cs, es);
    // refuse the whole transform before the first launch if any pass's grid is too large
    for (int p = 0; p < n_passes; p++) {
        const int kernel = ntt_pass_kernel(tu, log_d, starts[p], ks[p], cs[p], es[p], inverse);
        if (ntt_pass_workitems(kernel, log_d, starts[p], ks[p], cs[p], es[p], device_cus()) >> 32)
            return hipErrorInvalidConfiguration;
    }
    for (int q = 0; q < n_passes; q++) {
        const int p = inverse ? (n_passes - 1 - q) : q;
        NttPassArgs a;
        a.data = data;
        a.W = dd.W;
        a.X = dd.X;
        a.x_layers = dd.x_layers;
        a.msg = (q == 0 && !inverse) ? msg : nullptr;
        a.m = msg_log;
        a.log_d = log_d;
        a.layer0 = starts[p];
        a.k = ks[p];
        a.c = cs[p];
        a.e = es[p];
        a.tab_min = tu.tab_min;
        const uint64_t n_tiles = 1ull << (log_d - a.k - a.c - a.e);
        const int kernel = ntt_pass_kernel(tu, log_d, a.layer0, a.k, a.c, a.e, inverse);
        if (kernel == B64_NTT_KERNEL_R16) {
            launch_ntt_r16(a, s);
        } else if (kernel == B64_NTT_KERNEL_R16W8) {
            launch_ntt_r16w8(a, s);
        } else if (kernel == B64_NTT_KERNEL_T8) {
            launch_ntt_t8(a, (unsigned)n_tiles, s);
        } else if (kernel == B64_NTT_KERNEL_T4U) {
            launch_ntt_t4u(a, (unsigned)n_tiles, s);
        } else if (kernel != B64_NTT_KERNEL_GENERAL) {
            if (inverse)
                launch_ntt_tab<true>(a, (unsigned)n_tiles, kernel, tu, s);
            else
                launch_ntt_tab<false>(a, (unsigned)n_tiles, kernel, tu, s);
        } else {
            if (inverse)
                k_ntt_pass<true><<<(unsigned)n_tiles, BLOCK, 0, s>>>(a);
            else
                k_ntt_pass<false><<<(unsigned)n_tiles, BLOCK, 0, s>>>(a);
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// ---------------------------------------------------------------------------
// Merkle leaves: SHA-256 over the LE16 bytes of `batch` consecutive elements
// (hash/parallel_digest.rs:156-172).  Digest bytes are stored as 8 LE words.
// Every stage below has a per-block core, a single-job kernel and a multi-job
// kernel (block -> job through a prefix table) used by the batched path.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void load_block_words(uint32_t w[16], const uint4 *p) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const uint4 e = p[q];
        w[4 * q + 0] = bswap32(e.x);
        w[4 * q + 1] = bswap32(e.y);
        w[4 * q + 2] = bswap32(e.z);
        w[4 * q + 3] = bswap32(e.w);
    }
}

__device__ __forceinline__ void store_digest_be(uint4 *d, const uint32_t st[8]) {
    d[0] = make_uint4(bswap32(st[0]), bswap32(st[1]), bswap32(st[2]), bswap32(st[3]));
    d[1] = make_uint4(bswap32(st[4]), bswap32(st[5]), bswap32(st[6]), bswap32(st[7]));
}

__device__ __forceinline__ void leaf_hash_core(const LeafArgs &a, uint64_t blk) {
    const uint64_t leaf = blk * BLOCK + threadIdx.x;
    if (leaf >= a.n_leaves) return;
    const uint64_t batch = a.batch;
    const uint4 *p = a.elems + leaf * batch;
    uint32_t st[8];
    sha256_init(st);
    uint32_t w[16];
    const uint64_t n_full = batch >> 2;
    for (uint64_t b = 0; b < n_full; b++) {
        load_block_words(w, p + 4 * b);
        sha256_compress(st, w);
    }
    const int rem = (int)(batch & 3);
#pragma unroll
    for (int q = 0; q < 16; q++) w[q] = 0;
    for (int q = 0; q < rem; q++) {
        const uint4 e = p[n_full * 4 + q];
        w[4 * q + 0] = bswap32(e.x);
        w[4 * q + 1] = bswap32(e.y);
        w[4 * q + 2] = bswap32(e.z);
        w[4 * q + 3] = bswap32(e.w);
    }
    w[4 * rem] = 0x80000000u;
    const uint64_t bits = batch * 128ull;
    w[14] = (uint32_t)(bits >> 32);
    w[15] = (uint32_t)bits;
    sha256_compress(st, w);
    store_digest_be(a.digests + 2 * leaf, st);
}

__global__ __launch_bounds__(BLOCK) void k_leaf_hash(LeafArgs a) { leaf_hash_core(a, blockIdx.x); }

// Few long leaves (the cosets of a small FRI oracle, e.g. the last one's 2
// leaves of 128 elements): each leaf is a chain of 9..33 compressions, which
// one lane of an otherwise idle wave runs at ~4 clocks per instruction.  So one
// 64-lane block per leaf: its lanes first put K[i] + W[i] of every block of the
// leaf (data blocks and the padding block) in LDS -- the message schedules are
// independent across blocks -- then lane 0 runs only the rounds, with the
// schedule off its chain (14 VALU per round instead of ~22).
constexpr int LONG_LEAF_MAX_BLOCKS = 33;  // 128 elements + the padding block

bool leaf_hash_long_fits(uint64_t batch, uint64_t n_leaves) {
    return batch % 4 == 0 && batch >= 32 && batch / 4 + 1 <= LONG_LEAF_MAX_BLOCKS && n_leaves <= 256;
}

__device__ __forceinline__ void leaf_hash_long_core(const LeafArgs &a, uint64_t leaf,
                                                    uint32_t (*__restrict__ kw)[64]) {
    const int nb = (int)(a.batch >> 2);  // data blocks; block nb is the padding block
    const uint4 *p = a.elems + leaf * a.batch;
    for (int b = threadIdx.x; b <= nb; b += 64) {
        uint32_t w[16];
        if (b < nb) {
            load_block_words(w, p + 4 * b);
        } else {
#pragma unroll
            for (int q = 0; q < 16; q++) w[q] = 0;
            w[0] = 0x80000000u;
            const uint64_t bits = a.batch * 128ull;
            w[14] = (uint32_t)(bits >> 32);
            w[15] = (uint32_t)bits;
        }
#pragma unroll
        for (int i = 0; i < 64; i++) {
            uint32_t wi;
            if (i < 16) {
                wi = w[i];
            } else {
                const uint32_t w15 = w[(i - 15) & 15], w2 = w[(i - 2) & 15];
                const uint32_t s0 = xor3(rotr(w15, 7), rotr(w15, 18), w15 >> 3);
                const uint32_t s1 = xor3(rotr(w2, 17), rotr(w2, 19), w2 >> 10);
                wi = w[i & 15] + s0 + w[(i - 7) & 15] + s1;
                w[i & 15] = wi;
            }
            kw[b][i] = wi + SHA_K[i];
        }
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    uint32_t st[8];
    sha256_init(st);
    // each block's 64 words come into registers while the previous block's
    // rounds run (an LDS read per round would put its latency on the chain)
    uint4 cur[16], nxt[16];
    const uint4 *k4 = reinterpret_cast<const uint4 *>(&kw[0][0]);
#pragma unroll
    for (int q = 0; q < 16; q++) cur[q] = k4[q];
    for (int b = 0; b <= nb; b++) {
        const int bn = b < nb ? b + 1 : b;  // the last block re-reads itself (unused)
#pragma unroll
        for (int q = 0; q < 16; q++) nxt[q] = k4[16 * bn + q];
        sha256_rounds_kw(st, cur);
#pragma unroll
        for (int q = 0; q < 16; q++) cur[q] = nxt[q];
    }
    store_digest_be(a.digests + 2 * leaf, st);
}

__global__ __launch_bounds__(64) void k_leaf_hash_long(LeafArgs a) {
    __shared__ __attribute__((aligned(16))) uint32_t kw[LONG_LEAF_MAX_BLOCKS][64];
    leaf_hash_long_core(a, blockIdx.x, kw);
}

// Many jobs in one launch: block b belongs to the job j with
// start[j] <= b < start[j+1] and works on that job's block b - start[j].
__device__ __forceinline__ int job_of_block2(const uint32_t *__restrict__ start, int n_jobs, uint32_t b) {
    int lo = 0, hi = n_jobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(BLOCK) void k_leaf_hash_multi(const LeafArgs *__restrict__ jobs,
                                                           const uint32_t *__restrict__ start, int n_jobs) {
    const int j = job_of_block2(start, n_jobs, blockIdx.x);
    leaf_hash_core(jobs[j], blockIdx.x - start[j]);
}

// One block per leaf over many jobs (every job's leaves pass leaf_hash_long_fits).
__global__ __launch_bounds__(64) void k_leaf_hash_long_multi(const LeafArgs *__restrict__ jobs,
                                                             const uint32_t *__restrict__ start, int n_jobs) {
    __shared__ __attribute__((aligned(16))) uint32_t kw[LONG_LEAF_MAX_BLOCKS][64];
    const int j = job_of_block2(start, n_jobs, blockIdx.x);
    leaf_hash_long_core(jobs[j], blockIdx.x - start[j], kw);
}

// Leaves of 16 elements (the interleaved batch 2^4 of the commit) hashed and,
// optionally, folded by fold_interleaved in the same pass over the codeword:
// the 16 elements of leaf i are exactly the 16 inputs of folded element i.
// The fold multiplies by the four challenges through LDS window tables.
__device__ __forceinline__ void leaf16_fold_core(const Leaf16Args &a, uint64_t blk, uint4 *__restrict__ tabs) {
    const bool do_fold = a.folded != nullptr;
    if (do_fold && threadIdx.x < 4 * WIN_ROWS)
        win_build_row(tabs + (threadIdx.x >> 5) * WIN_ENTRIES, a.chal[threadIdx.x >> 5], threadIdx.x & 31);
    __syncthreads();
    const uint64_t leaf = blk * BLOCK + threadIdx.x;
    if (leaf >= a.n_leaves) return;
    const uint4 *p = a.cw + leaf * 16;
    uint4 v[16];
#pragma unroll
    for (int j = 0; j < 16; j++) v[j] = p[j];
    uint32_t st[8];
    sha256_init(st);
    uint32_t w[16];
#pragma unroll
    for (int b = 0; b < 4; b++) {
        load_block_words(w, v + 4 * b);
        sha256_compress(st, w);
    }
#pragma unroll
    for (int q = 0; q < 16; q++) w[q] = 0;
    w[0] = 0x80000000u;
    w[15] = 16 * 128;
    sha256_compress(st, w);
    store_digest_be(a.digests + 2 * leaf, st);
    if (do_fold) {
        // permute-addressed lookups: one VALU op per lookup address (the SHA rounds
        // keep this kernel VALU-bound)
        const uint32_t tb = lds_addr(tabs);
        int len = 16;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            len >>= 1;
#pragma unroll
            for (int q = 0; q < len; q++)
                v[q] = win_mul_perm(tb + (uint32_t)(t * WIN_ENTRIES * 16), gxor(v[2 * q], v[2 * q + 1]), v[2 * q]);
        }
        a.folded[leaf] = v[0];
    }
}

__global__ __launch_bounds__(BLOCK) void k_leaf16_fold(Leaf16Args a) {
    __shared__ __attribute__((aligned(256))) uint4 tabs[4 * WIN_ENTRIES];
    leaf16_fold_core(a, blockIdx.x, tabs);
}

__global__ __launch_bounds__(BLOCK) void k_leaf16_fold_multi(const Leaf16Args *__restrict__ jobs,
                                                             const uint32_t *__restrict__ start, int n_jobs) {
    __shared__ __attribute__((aligned(256))) uint4 tabs[4 * WIN_ENTRIES];
    const int j = job_of_block2(start, n_jobs, blockIdx.x);
    leaf16_fold_core(jobs[j], blockIdx.x - start[j], tabs);
}

hipError_t launch_leaf16_fold(const uint4 *cw, uint64_t n_leaves, uint4 *digests, uint4 *folded, const uint4 *chal,
                              hipStream_t s) {
    Leaf16Args a;
    a.cw = cw;
    a.n_leaves = n_leaves;
    a.digests = digests;
    a.folded = folded;
    for (int t = 0; t < 4; t++) a.chal[t] = folded ? chal[t] : make_uint4(0, 0, 0, 0);
    k_leaf16_fold<<<grid_for(n_leaves), BLOCK, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_leaf_hash(const uint4 *elems, uint64_t n_leaves, int, uint64_t batch, uint4 *digests,
                            hipStream_t s) {
    if (batch == 16) return launch_leaf16_fold(elems, n_leaves, digests, nullptr, nullptr, s);
    LeafArgs a;
    a.elems = elems;
    a.n_leaves = n_leaves;
    a.batch = batch;
    a.digests = digests;
    if (leaf_hash_long_fits(batch, n_leaves))
        k_leaf_hash_long<<<(unsigned)n_leaves, 64, 0, s>>>(a);
    else
        k_leaf_hash<<<grid_for(n_leaves), BLOCK, 0, s>>>(a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Inner nodes: Sha256Compression (hash/compress.rs:46-56): one compress256 of
// left||right from the custom IV, output state written little-endian.
// Subtree reduction: a block takes 2^levels consecutive nodes of one tree level
// (levels <= MERKLE_RED_LOG) and writes the next `levels` levels, keeping the
// working level in LDS.  The tree is stored level after level, so level l+1
// starts right after level l.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void node_compress(uint4 l0, uint4 l1, uint4 r0, uint4 r1, const CompressIV &iv,
                                              uint4 &o0, uint4 &o1) {
    uint32_t w[16] = {bswap32(l0.x), bswap32(l0.y), bswap32(l0.z), bswap32(l0.w),
                      bswap32(l1.x), bswap32(l1.y), bswap32(l1.z), bswap32(l1.w),
                      bswap32(r0.x), bswap32(r0.y), bswap32(r0.z), bswap32(r0.w),
                      bswap32(r1.x), bswap32(r1.y), bswap32(r1.z), bswap32(r1.w)};
    uint32_t st[8];
#pragma unroll
    for (int q = 0; q < 8; q++) st[q] = iv.w[q];
    sha256_compress(st, w);
    o0 = make_uint4(st[0], st[1], st[2], st[3]);
    o1 = make_uint4(st[4], st[5], st[6], st[7]);
}

// Block b reduces nodes [b*chunk, (b+1)*chunk) of the input level by
// a.levels levels, writing every intermediate level.
__device__ __forceinline__ void subtree_core(const SubtreeArgs &a, uint64_t b, uint4 (*dig)[2 << MERKLE_RED_LOG]) {
    const int blk_nodes = (int)(a.n_in < (1ull << MERKLE_RED_LOG) ? a.n_in : (1ull << MERKLE_RED_LOG));
    const uint4 *src = a.tree + 2 * (a.in_off + b * (uint64_t)blk_nodes);
    constexpr int PER = (2 << MERKLE_RED_LOG) / BLOCK;
    uint4 r[PER];
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int t = threadIdx.x + i * BLOCK;
        r[i] = t < 2 * blk_nodes ? src[t] : make_uint4(0, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int t = threadIdx.x + i * BLOCK;
        if (t < 2 * blk_nodes) dig[0][t] = r[i];
    }
    __syncthreads();
    uint64_t off = a.in_off, n = a.n_in;
    int cur = 0;
    for (int l = 1; l <= a.levels; l++) {
        const uint64_t out_off = off + n;
        const int cnt = blk_nodes >> l;
        uint4 *dst = a.tree + 2 * (out_off + b * (uint64_t)cnt);
        for (int t = threadIdx.x; t < cnt; t += BLOCK) {
            uint4 o0, o1;
            node_compress(dig[cur][4 * t], dig[cur][4 * t + 1], dig[cur][4 * t + 2], dig[cur][4 * t + 3], a.iv, o0,
                          o1);
            dig[cur ^ 1][2 * t] = o0;
            dig[cur ^ 1][2 * t + 1] = o1;
            dst[2 * t] = o0;
            dst[2 * t + 1] = o1;
        }
        __syncthreads();
        cur ^= 1;
        off = out_off;
        n >>= 1;
    }
}

__global__ __launch_bounds__(BLOCK) void k_merkle_subtree(SubtreeArgs a) {
    __shared__ uint4 dig[2][2 << MERKLE_RED_LOG];
    subtree_core(a, blockIdx.x, dig);
}

__global__ __launch_bounds__(BLOCK) void k_merkle_subtree_multi(const SubtreeArgs *__restrict__ jobs,
                                                                const uint32_t *__restrict__ start, int n_jobs) {
    __shared__ uint4 dig[2][2 << MERKLE_RED_LOG];
    const int j = job_of_block2(start, n_jobs, blockIdx.x);
    subtree_core(jobs[j], blockIdx.x - start[j], dig);
}

// Two levels per thread: thread t compresses input nodes 4t..4t+3 into 2 nodes
// of the next level and those into 1 node of the level after.  The two first
// compressions are independent, so a lane keeps two chains in flight.
__device__ __forceinline__ void merkle_quad_core(const SubtreeArgs &a, uint64_t t) {
    if (t >= (a.n_in >> 2)) return;
    const uint4 *src = a.tree + 2 * (a.in_off + 4 * t);
    uint4 x[8];
#pragma unroll
    for (int i = 0; i < 8; i++) x[i] = src[i];
    uint4 p0, p1, q0, q1, r0, r1;
    node_compress(x[0], x[1], x[2], x[3], a.iv, p0, p1);
    node_compress(x[4], x[5], x[6], x[7], a.iv, q0, q1);
    uint4 *l1 = a.tree + 2 * (a.in_off + a.n_in + 2 * t);
    l1[0] = p0;
    l1[1] = p1;
    l1[2] = q0;
    l1[3] = q1;
    node_compress(p0, p1, q0, q1, a.iv, r0, r1);
    uint4 *l2 = a.tree + 2 * (a.in_off + a.n_in + (a.n_in >> 1) + t);
    l2[0] = r0;
    l2[1] = r1;
}

__global__ __launch_bounds__(BLOCK) void k_merkle_quad(SubtreeArgs a) {
    merkle_quad_core(a, (uint64_t)blockIdx.x * BLOCK + threadIdx.x);
}

__global__ __launch_bounds__(BLOCK) void k_merkle_quad_multi(const SubtreeArgs *__restrict__ jobs,
                                                             const uint32_t *__restrict__ start, int n_jobs) {
    const int j = job_of_block2(start, n_jobs, blockIdx.x);
    merkle_quad_core(jobs[j], (uint64_t)(blockIdx.x - start[j]) * BLOCK + threadIdx.x);
}

uint32_t merkle_step_blocks(uint64_t n_in, int quad) {
    return quad ? (uint32_t)(((n_in >> 2) + BLOCK - 1) / BLOCK) : (uint32_t)(n_in / merkle_chunk(n_in));
}

int merkle_reduce_steps(uint64_t n_leaves, uint64_t *offs, uint64_t *ns, int *levels_out, int *quad) {
    uint64_t off = 0, n = n_leaves;
    int steps = 0;
    while (n > 1) {
        int log_n = 0;
        while ((1ull << log_n) < n) log_n++;
        const bool q = log_n >= MERKLE_QUAD_LOG;
        const int levels = q ? 2
                             : log_n > MERKLE_WIDE_LOG ? MERKLE_WIDE_LEVELS
                                                       : (log_n < MERKLE_RED_LOG ? log_n : MERKLE_RED_LOG);
        if (offs) {
            offs[steps] = off;
            ns[steps] = n;
            levels_out[steps] = levels;
            quad[steps] = q ? 1 : 0;
        }
        steps++;
        for (int l = 0; l < levels; l++) {
            off += n;
            n >>= 1;
        }
    }
    return steps;
}

hipError_t launch_merkle_levels(uint4 *tree, uint64_t n_leaves, const CompressIV &iv, hipStream_t s) {
    uint64_t offs[32], ns[32];
    int lv[32], qd[32];
    if (merkle_reduce_steps(n_leaves, nullptr, nullptr, nullptr, nullptr) > 32) return hipErrorInvalidValue;
    const int steps = merkle_reduce_steps(n_leaves, offs, ns, lv, qd);
    for (int i = 0; i < steps; i++) {
        SubtreeArgs a;
        a.tree = tree;
        a.in_off = offs[i];
        a.n_in = ns[i];
        a.levels = lv[i];
        a.iv = iv;
        if (qd[i])
            k_merkle_quad<<<merkle_step_blocks(ns[i], 1), BLOCK, 0, s>>>(a);
        else
            k_merkle_subtree<<<merkle_step_blocks(ns[i], 0), BLOCK, 0, s>>>(a);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// ---------------------------------------------------------------------------
// FRI folds
// ---------------------------------------------------------------------------
struct FoldChallenges {
    uint4 r[MAX_FOLD_CHAL];
};

// fold_interleaved (prover fri/fold.rs:260-282): out[i] = <cw[i*B..(i+1)*B], eq(r)>.
// Computed as log_batch successive line folds (bit t of the in-row index pairs
// with r_t, the eq tensor's bit order, multilinear/eq.rs:39-71); the result is
// the same field element as the inner product.
template <int LOGB>
__device__ __forceinline__ void fold_interleaved_core(const uint4 *__restrict__ cw, uint4 *__restrict__ out,
                                                      uint64_t n_out, const uint4 *chal, uint64_t blk,
                                                      uint4 *__restrict__ tabs) {
    for (int r = threadIdx.x; r < LOGB * WIN_ROWS; r += BLOCK)
        win_build_row(tabs + (r >> 5) * WIN_ENTRIES, chal[r >> 5], r & 31);
    __syncthreads();
    const uint64_t i = blk * BLOCK + threadIdx.x;
    if (i >= n_out) return;
    constexpr int B = 1 << LOGB;
    uint4 v[B];
#pragma unroll
    for (int j = 0; j < B; j++) v[j] = cw[i * B + j];
    int len = B;
#pragma unroll
    for (int t = 0; t < LOGB; t++) {
        len >>= 1;
#pragma unroll
        for (int q = 0; q < len; q++) v[q] = gxor(v[2 * q], win_mul(tabs + t * WIN_ENTRIES, gxor(v[2 * q], v[2 * q + 1])));
    }
    out[i] = v[0];
}

template <int LOGB>
__global__ __launch_bounds__(BLOCK) void k_fold_interleaved(const uint4 *__restrict__ cw, uint4 *__restrict__ out,
                                                            uint64_t n_out, FoldChallenges ch) {
    __shared__ uint4 tabs[(LOGB > 0 ? LOGB : 1) * WIN_ENTRIES];
    fold_interleaved_core<LOGB>(cw, out, n_out, ch.r, blockIdx.x, tabs);
}

hipError_t launch_fold_interleaved(const uint4 *cw, uint4 *out, uint64_t n_out, int log_batch,
                                   const uint4 *chal, hipStream_t s) {
    FoldChallenges ch;
    for (int t = 0; t < log_batch; t++) ch.r[t] = chal[t];
    const unsigned g = grid_for(n_out);
    switch (log_batch) {
        case 0: k_fold_interleaved<0><<<g, BLOCK, 0, s>>>(cw, out, n_out, ch); break;
        case 1: k_fold_interleaved<1><<<g, BLOCK, 0, s>>>(cw, out, n_out, ch); break;
        case 2: k_fold_interleaved<2><<<g, BLOCK, 0, s>>>(cw, out, n_out, ch); break;
        case 3: k_fold_interleaved<3><<<g, BLOCK, 0, s>>>(cw, out, n_out, ch); break;
        case 4: k_fold_interleaved<4><<<g, BLOCK, 0, s>>>(cw, out, n_out, ch); break;
        case 5: k_fold_interleaved<5><<<g, BLOCK, 0, s>>>(cw, out, n_out, ch); break;
        case 6: k_fold_interleaved<6><<<g, BLOCK, 0, s>>>(cw, out, n_out, ch); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// twiddle(layer, block) from the layer's F2 basis.
__device__ __forceinline__ uint4 twiddle_of(const uint4 *__restrict__ W, int layer, uint64_t block) {
    const uint4 *Wl = W + layer * W_STRIDE;
    uint4 t = make_uint4(0, 0, 0, 0);
    for (int b = 0; b < layer; b++)
        if ((block >> b) & 1) t = gxor(t, Wl[b]);
    return t;
}

// fold_codeword (prover fri/fold.rs:297-318) with fold_chunk / fold_pair
// (verifier fri/fold.rs:21-99): per coset, n_chal rounds of inverse butterfly
// and line extrapolation x0 + (x1 - x0) * r.  General form for n_chal 5..8
// (the coset's values do not fit in registers).
template <int MAXV>
__device__ __forceinline__ void fold_codeword_core(const FoldCwArgs &a, uint64_t blk) {
    const uint64_t i = blk * BLOCK + threadIdx.x;
    if (i >= a.n_out) return;
    const int n_chal = a.n_chal;
    uint4 v[MAXV];
    const int cs = 1 << n_chal;
    for (int j = 0; j < cs; j++) v[j] = a.cw[(i << n_chal) + j];
    int ll = a.log_len, log_size = n_chal;
    for (int r = 0; r < n_chal; r++) {
        for (int off = 0; off < (1 << (log_size - 1)); off++) {
            const uint64_t idx = (i << (log_size - 1)) | (uint64_t)off;
            const uint4 t = twiddle_of(a.W, ll - 1, idx);
            uint4 u = v[off << 1], w = v[(off << 1) | 1];
            w = gxor(w, u);
            u = gxor(u, gf_mul(w, t));
            v[off] = gxor(u, gf_mul(gxor(w, u), a.chal[r]));
        }
        ll--;
        log_size--;
    }
    a.out[i] = v[0];
}

// The same fold for n_chal = NC <= 4, with the coset in registers.  Round r
// works at layer log_len - 1 - r on blocks (i << s) | off, s = NC - 1 - r: the
// twiddle is the XOR of the layer's basis elements selected by the block's bits
// (twiddle_of), i.e. one sum over i's bits shared by the round plus, per off,
// the basis elements below s.  Rounds and butterflies are unrolled by template
// recursion: as loops they are too large for the unroller, and the coset array
// then lands in scratch memory.  (Window tables for the challenge products were
// tried: next to the general product in the same chain the compiler runs out
// of registers.)
template <int S, int OFF>
__device__ __forceinline__ void fold_bfly_reg(uint4 *v, uint4 base, const uint4 *__restrict__ Wl, uint4 ch) {
    if constexpr (OFF < (1 << S)) {
        uint4 t = base;
#pragma unroll
        for (int b = 0; b < S; b++)
            if ((OFF >> b) & 1) t = gxor(t, Wl[b]);
        uint4 u = v[OFF << 1], w = v[(OFF << 1) | 1];
        w = gxor(w, u);
        u = gxor(u, gf_mul(w, t));
        v[OFF] = gxor(u, gf_mul(gxor(w, u), ch));
        fold_bfly_reg<S, OFF + 1>(v, base, Wl, ch);
    }
}

template <int NC, int R>
__device__ __forceinline__ void fold_rounds_reg(uint4 *v, const uint4 *__restrict__ W, int log_len, const uint4 *ch,
                                                uint64_t i) {
    if constexpr (R < NC) {
        constexpr int S = NC - 1 - R;
        const int layer = log_len - 1 - R;  // >= S, since log_len >= NC
        const uint4 *Wl = W + layer * W_STRIDE;
        uint4 base = make_uint4(0, 0, 0, 0);
        for (int b = S; b < layer; b++) {
            const uint32_t m = 0u - (uint32_t)((i >> (b - S)) & 1);
            const uint4 wb = Wl[b];
            base = make_uint4(base.x ^ (wb.x & m), base.y ^ (wb.y & m), base.z ^ (wb.z & m), base.w ^ (wb.w & m));
        }
        fold_bfly_reg<S, 0>(v, base, Wl, ch[R]);
        fold_rounds_reg<NC, R + 1>(v, W, log_len, ch, i);
    }
}

__device__ __forceinline__ uint32_t sreg(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ uint4 sreg4(uint4 v) { return make_uint4(sreg(v.x), sreg(v.y), sreg(v.z), sreg(v.w)); }
template <typename T>
__device__ __forceinline__ T *sreg_ptr(T *p) {
    const uint64_t x = (uint64_t)p;
    return (T *)(((uint64_t)sreg((uint32_t)(x >> 32)) << 32) | sreg((uint32_t)x));
}

// A job's fold, its arguments read from device memory (the same for the whole
// block, so they are moved to scalar registers).
template <int NC>
__device__ __forceinline__ void fold_codeword_job(const FoldCwArgs &a, uint64_t i) {
    uint4 ch[NC];
#pragma unroll
    for (int r = 0; r < NC; r++) ch[r] = sreg4(a.chal[r]);
    const uint4 *cw = sreg_ptr(a.cw), *W = sreg_ptr(a.W);
    uint4 *out = sreg_ptr(a.out);
    const int log_len = (int)sreg((uint32_t)a.log_len);
    uint4 v[1 << NC];
#pragma unroll
    for (int j = 0; j < (1 << NC); j++) v[j] = cw[(i << NC) + j];
    fold_rounds_reg<NC, 0>(v, W, log_len, ch, i);
    out[i] = v[0];
}

// n_chal 0..4 (the batch plans reject larger arities).
__device__ __forceinline__ void fold_codeword_upto4(const FoldCwArgs &a, uint64_t blk) {
    const uint64_t i = blk * BLOCK + threadIdx.x;
    if (i >= a.n_out) return;
    switch (a.n_chal) {
        case 0: a.out[i] = a.cw[i]; break;
        case 1: fold_codeword_job<1>(a, i); break;
        case 2: fold_codeword_job<2>(a, i); break;
        case 3: fold_codeword_job<3>(a, i); break;
        case 4: fold_codeword_job<4>(a, i); break;
        default: break;
    }
}

template <int NC>
__global__ __launch_bounds__(BLOCK) void k_fold_codeword_reg(FoldCwArgs a) {
    const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= a.n_out) return;
    uint4 ch[NC];
#pragma unroll
    for (int r = 0; r < NC; r++) ch[r] = a.chal[r];
    uint4 v[1 << NC];
#pragma unroll
    for (int j = 0; j < (1 << NC); j++) v[j] = a.cw[(i << NC) + j];
    fold_rounds_reg<NC, 0>(v, a.W, a.log_len, ch, i);
    a.out[i] = v[0];
}

template <int MAXV>
__global__ __launch_bounds__(BLOCK) void k_fold_codeword(FoldCwArgs a) {
    fold_codeword_core<MAXV>(a, blockIdx.x);
}

// Jobs of mixed arity (the switch costs registers: one wave per SIMD).
__global__ __launch_bounds__(BLOCK) void k_fold_codeword_multi(const FoldCwArgs *__restrict__ jobs,
                                                               const uint32_t *__restrict__ start, int n_jobs) {
    const int j = job_of_block2(start, n_jobs, blockIdx.x);
    fold_codeword_upto4(jobs[j], blockIdx.x - start[j]);
}

// Jobs that all fold NC challenges.
template <int NC>
__global__ __launch_bounds__(BLOCK) void k_fold_codeword_multi_nc(const FoldCwArgs *__restrict__ jobs,
                                                                  const uint32_t *__restrict__ start, int n_jobs) {
    // the job index is block-uniform: say so, so that the job's fields load
    // into scalar registers
    const int j = __builtin_amdgcn_readfirstlane(job_of_block2(start, n_jobs, blockIdx.x));
    const FoldCwArgs &a = jobs[j];
    const uint64_t i = (uint64_t)(blockIdx.x - start[j]) * BLOCK + threadIdx.x;
    if (i < a.n_out) fold_codeword_job<NC>(a, i);
}

hipError_t launch_fold_codeword(const DeviceDomain &dd, const uint4 *cw, uint4 *out, int log_len,
                                const uint4 *chal, int n_chal, hipStream_t s) {
    FoldCwArgs a;
    a.cw = cw;
    a.out = out;
    a.n_out = 1ull << (log_len - n_chal);
    a.log_len = log_len;
    a.n_chal = n_chal;
    a.W = dd.W;
    if (n_chal > MAX_FOLD_CHAL) return hipErrorInvalidValue;
    for (int t = 0; t < n_chal; t++) a.chal[t] = chal[t];
    if (n_chal == 0)
        k_fold_codeword<16><<<grid_for(a.n_out), BLOCK, 0, s>>>(a);
    else if (n_chal == 1)
        k_fold_codeword_reg<1><<<grid_for(a.n_out), BLOCK, 0, s>>>(a);
    else if (n_chal == 2)
        k_fold_codeword_reg<2><<<grid_for(a.n_out), BLOCK, 0, s>>>(a);
    else if (n_chal == 3)
        k_fold_codeword_reg<3><<<grid_for(a.n_out), BLOCK, 0, s>>>(a);
    else if (n_chal == 4)
        k_fold_codeword_reg<4><<<grid_for(a.n_out), BLOCK, 0, s>>>(a);
    else if (n_chal <= 8)
        k_fold_codeword<256><<<grid_for(a.n_out), BLOCK, 0, s>>>(a);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

template <int G>
__global__ __launch_bounds__(BLOCK) void k_ntt_pass_tab_multi(const NttPassArgs *__restrict__ jobs,
                                                              const uint32_t *__restrict__ start, int n_jobs) {
    __shared__ NttTabSmem<G> sm;
    const int j = job_of_block2(start, n_jobs, blockIdx.x);
    const NttPassArgs a = jobs[j];
    ntt_tab_tile<false, G, true>(a, blockIdx.x - start[j], sm);
}

// fold_interleaved (batch 2^4) over many jobs.
__global__ __launch_bounds__(BLOCK) void k_fold_interleaved_multi(const FoldIlArgs *__restrict__ jobs,
                                                                 const uint32_t *__restrict__ start, int n_jobs) {
    __shared__ uint4 tabs[4 * WIN_ENTRIES];
    const int j = job_of_block2(start, n_jobs, blockIdx.x);
    const FoldIlArgs &a = jobs[j];
    fold_interleaved_core<4>(a.cw, a.out, a.n_out, a.chal, blockIdx.x - start[j], tabs);
}

// The register and small-table passes over many jobs: each job brings its own
// pass arguments and positions per block (NttJobArgs).
template <bool MSG>
__global__ __launch_bounds__(R16_BLOCK) void k_ntt_pass_r16_multi(const NttJobArgs *__restrict__ jobs,
                                                                 const uint32_t *__restrict__ start, int n_jobs) {
    __shared__ R16Smem sm;
    const int j = job_of_block2(start, n_jobs, blockIdx.x);
    const NttJobArgs ja = jobs[j];
    ntt_r16_core<MSG>(ja.a, ja.per_block, blockIdx.x - start[j], sm);
}

template <bool MSG>
__global__ __launch_bounds__(R8_BLOCK) void k_ntt_pass_r16w8_multi(const NttJobArgs *__restrict__ jobs,
                                                                  const uint32_t *__restrict__ start, int n_jobs) {
    __shared__ R8Smem sm;
    const int j = job_of_block2(start, n_jobs, blockIdx.x);
    const NttJobArgs ja = jobs[j];
    ntt_r16w8_core<MSG>(ja.a, ja.per_block, blockIdx.x - start[j], sm);
}

template <int TL>
__global__ __launch_bounds__(BLOCK) void k_ntt_pass_t4u_multi(const NttJobArgs *__restrict__ jobs,
                                                             const uint32_t *__restrict__ start, int n_jobs) {
    __shared__ NttT4USmem<TL> sm;
    const int j = job_of_block2(start, n_jobs, blockIdx.x);
    const NttJobArgs ja = jobs[j];
    ntt_t4u_core<TL>(ja.a, blockIdx.x - start[j], sm);
}

// Multi-job kind, positions per block and block count of one job's pass under a
// batch schedule: with batch variant >= 3 the pass runs the kernel a single-job
// launch would (register passes, the one-round small-table tile), in its
// multi-job form; other passes, and variants 0-1, run the multi-job table kernel.
int ntt_multi_pass(const NttTuning &tu, int log_d, int layer0, int k, int c, int e, bool msg, MultiKind *kind,
                   uint64_t *per_block, uint32_t *blocks) {
    *per_block = 0;
    if (tu.variant >= 3) {
        const int kernel = ntt_pass_kernel(tu, log_d, layer0, k, c, e, false);
        const uint32_t cus = device_cus();
        if (kernel == B64_NTT_KERNEL_R16) {
            const int per_log = r16_per_log(log_d, layer0, cus);
            *kind = msg ? MultiKind::NttR16Msg : MultiKind::NttR16;
            *per_block = 1ull << per_log;
            *blocks = 1u << (log_d - 4 - per_log);
            return B64_NTT_KERNEL_R16;
        }
        if (kernel == B64_NTT_KERNEL_R16W8) {
            const int per_log = r8_per_log(log_d, layer0, cus);
            *kind = msg ? MultiKind::NttR16W8Msg : MultiKind::NttR16W8;
            *per_block = 1ull << per_log;
            *blocks = 1u << (log_d - 4 - per_log);
            return B64_NTT_KERNEL_R16W8;
        }
        if (kernel == B64_NTT_KERNEL_T4U) {
            *kind = k + c + e <= 9 ? MultiKind::NttT4U9 : MultiKind::NttT4U10;
            *blocks = 1u << (log_d - k - c - e);
            return B64_NTT_KERNEL_T4U;
        }
    }
    *kind = MultiKind::NttPass;
    *blocks = 1u << (log_d - k - c - e);
    return B64_NTT_KERNEL_TAB_HYBRID;
}

// ---------------------------------------------------------------------------
// Multi-job launchers (batched path): one launch covers the same stage of many
// jobs; tables live in device memory prepared by the caller.
// ---------------------------------------------------------------------------
hipError_t launch_multi(MultiKind kind, const void *d_jobs, const uint32_t *d_start, int n_jobs, uint32_t blocks,
                        hipStream_t s, int fold_nc) {
    if (blocks == 0 || n_jobs == 0) return hipSuccess;
    switch (kind) {
        case MultiKind::NttPass:
            switch (g_batch_tuning.tab_group) {
                case 8:
                    k_ntt_pass_tab_multi<8><<<blocks, BLOCK, 0, s>>>((const NttPassArgs *)d_jobs, d_start, n_jobs);
                    break;
                case 4:
                    k_ntt_pass_tab_multi<4><<<blocks, BLOCK, 0, s>>>((const NttPassArgs *)d_jobs, d_start, n_jobs);
                    break;
                default:
                    k_ntt_pass_tab_multi<2><<<blocks, BLOCK, 0, s>>>((const NttPassArgs *)d_jobs, d_start, n_jobs);
                    break;
            }
            break;
        case MultiKind::NttR16:
            k_ntt_pass_r16_multi<false><<<blocks, R16_BLOCK, 0, s>>>((const NttJobArgs *)d_jobs, d_start, n_jobs);
            break;
        case MultiKind::NttR16Msg:
            k_ntt_pass_r16_multi<true><<<blocks, R16_BLOCK, 0, s>>>((const NttJobArgs *)d_jobs, d_start, n_jobs);
            break;
        case MultiKind::NttR16W8:
            k_ntt_pass_r16w8_multi<false><<<blocks, R8_BLOCK, 0, s>>>((const NttJobArgs *)d_jobs, d_start, n_jobs);
            break;
        case MultiKind::NttR16W8Msg:
            k_ntt_pass_r16w8_multi<true><<<blocks, R8_BLOCK, 0, s>>>((const NttJobArgs *)d_jobs, d_start, n_jobs);
            break;
        case MultiKind::NttT4U9:
            k_ntt_pass_t4u_multi<9><<<blocks, BLOCK, 0, s>>>((const NttJobArgs *)d_jobs, d_start, n_jobs);
            break;
        case MultiKind::NttT4U10:
            k_ntt_pass_t4u_multi<10><<<blocks, BLOCK, 0, s>>>((const NttJobArgs *)d_jobs, d_start, n_jobs);
            break;
        case MultiKind::Leaf16Fold:
            k_leaf16_fold_multi<<<blocks, BLOCK, 0, s>>>((const Leaf16Args *)d_jobs, d_start, n_jobs);
            break;
        case MultiKind::LeafHash:
            k_leaf_hash_multi<<<blocks, BLOCK, 0, s>>>((const LeafArgs *)d_jobs, d_start, n_jobs);
            break;
        case MultiKind::LeafHashLong:
            k_leaf_hash_long_multi<<<blocks, 64, 0, s>>>((const LeafArgs *)d_jobs, d_start, n_jobs);
            break;
        case MultiKind::Subtree:
            k_merkle_subtree_multi<<<blocks, BLOCK, 0, s>>>((const SubtreeArgs *)d_jobs, d_start, n_jobs);
            break;
        case MultiKind::MerkleQuad:
            k_merkle_quad_multi<<<blocks, BLOCK, 0, s>>>((const SubtreeArgs *)d_jobs, d_start, n_jobs);
            break;
        case MultiKind::FoldCodeword: {
            const FoldCwArgs *j = (const FoldCwArgs *)d_jobs;
            switch (fold_nc) {
                case 1: k_fold_codeword_multi_nc<1><<<blocks, BLOCK, 0, s>>>(j, d_start, n_jobs); break;
                case 2: k_fold_codeword_multi_nc<2><<<blocks, BLOCK, 0, s>>>(j, d_start, n_jobs); break;
                case 3: k_fold_codeword_multi_nc<3><<<blocks, BLOCK, 0, s>>>(j, d_start, n_jobs); break;
                case 4: k_fold_codeword_multi_nc<4><<<blocks, BLOCK, 0, s>>>(j, d_start, n_jobs); break;
                default: k_fold_codeword_multi<<<blocks, BLOCK, 0, s>>>(j, d_start, n_jobs); break;
            }
            break;
        }
        case MultiKind::FoldInterleaved:
            k_fold_interleaved_multi<<<blocks, BLOCK, 0, s>>>((const FoldIlArgs *)d_jobs, d_start, n_jobs);
            break;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// pack_witness (prover/prover/src/prove.rs:286-327), zero padded.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_pack_witness(const uint64_t *__restrict__ words, uint64_t n_words,
                                                        uint4 *__restrict__ out, uint64_t n_elems) {
    const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n_elems) return;
    const uint64_t w0 = 2 * i < n_words ? words[2 * i] : 0;
    const uint64_t w1 = 2 * i + 1 < n_words ? words[2 * i + 1] : 0;
    out[i] = make_uint4((uint32_t)w0, (uint32_t)(w0 >> 32), (uint32_t)w1, (uint32_t)(w1 >> 32));
}

hipError_t launch_pack_witness(const uint64_t *words, uint64_t n_words, uint4 *out, uint64_t n_elems,
                               hipStream_t s) {
    k_pack_witness<<<grid_for(n_elems), BLOCK, 0, s>>>(words, n_words, out, n_elems);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// FRI query phase (prover fri/fold.rs:221-244, query.rs:51-134): one thread per
// 16-byte granule of a job's tape -- the terminal codeword, the optimal layer of
// every opened oracle, then per query index each oracle's coset values and
// branch digests (prove_coset_opening, query.rs:112-134; siblings as in
// BinaryMerkleTree::branch, binary_merkle_tree.rs:154-169).  The tree of an
// oracle stores its levels leaves first, so level l starts at node
// 2^(t+1) - 2^(t+1-l) and the layer of 2^depth digests at 2^(t+1) - 2^(depth+1).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_fri_query_multi(const QueryArgs *__restrict__ jobs,
                                                           const uint32_t *__restrict__ start, int n_jobs,
                                                           const uint64_t *__restrict__ idx, uint4 *__restrict__ out) {
    const int j = __builtin_amdgcn_readfirstlane(job_of_block2(start, n_jobs, blockIdx.x));
    const QueryArgs &a = jobs[j];
    uint64_t i = (uint64_t)(blockIdx.x - start[j]) * BLOCK + threadIdx.x;
    if (i >= a.n_gran) return;
    uint4 *dst = out + a.out_at + i;
    if (i < a.term_gran) {
        *dst = a.term[i];
        return;
    }
    i -= a.term_gran;
    if (i < a.layers_gran) {
        for (int k = 0; k < a.n_opened; k++) {
            const QueryOracle &o = a.o[k];
            const uint64_t n = 2ull << o.depth;  // 2^depth digests
            if (i < n) {
                *dst = o.tree[2 * ((2ull << o.tree_log_len) - n) + i];
                return;
            }
            i -= n;
        }
        return;
    }
    const uint32_t r = (uint32_t)(i - a.layers_gran);
    const uint32_t q = r / a.per_query;
    uint32_t rr = r - q * a.per_query;
    uint64_t cur = idx[a.idx_at + q];
    for (int k = 0; k < a.n_opened; k++) {
        const QueryOracle &o = a.o[k];
        if (k > 0) cur >>= o.log_coset;
        if (rr < o.n_gran) {
            cur &= (1ull << o.tree_log_len) - 1;  // the host has checked the range; this keeps reads in bounds
            const uint32_t nv = 1u << o.log_coset;
            if (rr < nv) {
                *dst = o.cw[(cur << o.log_coset) + rr];
                return;
            }
            const uint32_t b = rr - nv, lev = b >> 1;
            const uint64_t node = (((1ull << lev) - 1) << (o.tree_log_len + 1 - lev)) | ((cur >> lev) ^ 1);
            *dst = o.tree[2 * node + (b & 1)];
            return;
        }
        rr -= o.n_gran;
    }
}

uint32_t fri_query_blocks(const QueryArgs &a) { return (uint32_t)((a.n_gran + BLOCK - 1) / BLOCK); }

hipError_t launch_fri_query(const QueryArgs *d_jobs, const uint32_t *d_start, int n_jobs, uint32_t blocks,
                            const uint64_t *d_idx, uint4 *d_out, hipStream_t s) {
    if (n_jobs == 0 || blocks == 0) return hipSuccess;
    k_fri_query_multi<<<blocks, BLOCK, 0, s>>>(d_jobs, d_start, n_jobs, d_idx, d_out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Ring switch (prover/prover/src/ring_switch.rs)
// ---------------------------------------------------------------------------

// fold_b128_elems_inplace (ring_switch.rs:130-176): x <- sum_b bit_b(x) * vec[b].
// The map is F2-linear in x, so it is a window-table product: row k, entry n of
// the table holds sum_i n_i * vec[4k + i] (same swizzled layout as the twiddle
// tables), and win_mul_perm sums the 32 rows selected by x's nibbles.
__global__ __launch_bounds__(BLOCK) void k_rs_fold_elems(uint4 *__restrict__ elems, uint64_t n, RsVec v) {
    __shared__ __attribute__((aligned(256))) uint4 tab[WIN_ENTRIES];
    for (int e = threadIdx.x; e < WIN_ENTRIES; e += BLOCK) {
        const int k = e >> 4, nib = e & 15;
        uint4 acc = make_uint4(0, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; i++)
            if ((nib >> i) & 1) acc = gxor(acc, v.v[4 * k + i]);
        tab[k * 16 + (nib ^ (k & 15))] = acc;
    }
    __syncthreads();
    const uint32_t tb = lds_addr(tab);
    for (uint64_t x = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; x < n; x += (uint64_t)gridDim.x * BLOCK)
        elems[x] = win_mul_perm(tb, elems[x]);
}

hipError_t launch_rs_fold_elems(uint4 *elems, uint64_t n, const RsVec &v, hipStream_t s) {
    const uint64_t blocks = (n + BLOCK - 1) / BLOCK;
    k_rs_fold_elems<<<(unsigned)(blocks < 8192 ? blocks : 8192), BLOCK, 0, s>>>(elems, n, v);
    return hipGetLastError();
}

// fold_1b_rows_for_b128 (ring_switch.rs:197-265): out[b] = sum_r vec[r] * bit_b(mat[r]).
// Method of four Russians as in the reference: per chunk of 4 rows the 16 subset
// sums of the vec rows are shared by all 128 output bits, and output bit b adds
// the sum indexed by bit b of the 4 mat rows.  As in the reference, each chunk's
// mat rows are first transposed in 4x4 bit blocks (square_transpose_const_size,
// ring_switch.rs:267-320), so the index for bit b = 4q + j is nibble q of
// transposed row j: one LDS read per chunk and bit.
// A block takes batches of 256 rows (one mat and one vec row per lane; the next
// batch's loads are issued before the current batch is processed and stay in
// flight, since the block synchronises through LDS only).  Lanes transpose one
// 32-bit word of one chunk and build 4 subset sums each; lane pairs share an
// output bit and take alternate chunks.  Each block writes 128 partial sums;
// k_rs_reduce XORs them.
constexpr int RS_BATCH = BLOCK;

// In-place 4x4 transpose of bit blocks within every nibble of four words:
// afterwards bit 4q + r of word j is bit 4q + j of the original word r.
__device__ __forceinline__ void transpose4_nibbles(uint32_t &e0, uint32_t &e1, uint32_t &e2, uint32_t &e3) {
    uint32_t t;
    t = ((e0 >> 1) ^ e1) & 0x55555555u;
    e1 ^= t;
    e0 ^= t << 1;
    t = ((e2 >> 1) ^ e3) & 0x55555555u;
    e3 ^= t;
    e2 ^= t << 1;
    t = ((e0 >> 2) ^ e2) & 0x33333333u;
    e2 ^= t;
    e0 ^= t << 2;
    t = ((e1 >> 2) ^ e3) & 0x33333333u;
    e3 ^= t;
    e1 ^= t << 2;
}

__global__ __launch_bounds__(BLOCK) void k_rs_fold_1b_rows(const uint4 *__restrict__ mat,
                                                           const uint4 *__restrict__ vec, uint64_t n,
                                                           uint4 *__restrict__ partial) {
    __shared__ uint4 m_rows[RS_BATCH], v_rows[RS_BATCH];
    __shared__ uint4 sums[RS_BATCH / 4][16];
    __shared__ uint32_t mt[RS_BATCH / 4][4][4];  // [chunk][transposed row j][word]
    __shared__ uint4 other[128];
    const int b = threadIdx.x & 127, half = threadIdx.x >> 7;
    const int j = b & 3, q = b >> 2, qw = q >> 3, qs = 4 * (q & 7);
    const uint32_t *m_words = (const uint32_t *)m_rows;
    const uint64_t stride = (uint64_t)gridDim.x * RS_BATCH;
    uint4 acc = make_uint4(0, 0, 0, 0);
    uint64_t r0 = (uint64_t)blockIdx.x * RS_BATCH;
    uint4 m_next = make_uint4(0, 0, 0, 0), v_next = make_uint4(0, 0, 0, 0);
    if (r0 + threadIdx.x < n) {
        m_next = mat[r0 + threadIdx.x];
        v_next = vec[r0 + threadIdx.x];
    }
    for (; r0 < n; r0 += stride) {
        m_rows[threadIdx.x] = m_next;  // rows past n are zero and add nothing
        v_rows[threadIdx.x] = v_next;
        lds_barrier();
        const uint64_t rn = r0 + stride + threadIdx.x;
        m_next = rn < n ? mat[rn] : make_uint4(0, 0, 0, 0);
        v_next = rn < n ? vec[rn] : make_uint4(0, 0, 0, 0);
        {
            const int g = threadIdx.x >> 2, w = threadIdx.x & 3;
            uint32_t e0 = m_words[4 * (4 * g + 0) + w], e1 = m_words[4 * (4 * g + 1) + w];
            uint32_t e2 = m_words[4 * (4 * g + 2) + w], e3 = m_words[4 * (4 * g + 3) + w];
            transpose4_nibbles(e0, e1, e2, e3);
            mt[g][0][w] = e0;
            mt[g][1][w] = e1;
            mt[g][2][w] = e2;
            mt[g][3][w] = e3;
        }
        // expand_subset_sums_array (verifier/math/src/span.rs:43-57) of each chunk
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int e = threadIdx.x + k * BLOCK, g = e >> 4, m = e & 15;
            uint4 x = make_uint4(0, 0, 0, 0);
#pragma unroll
            for (int i = 0; i < 4; i++)
                if ((m >> i) & 1) x = gxor(x, v_rows[4 * g + i]);
            sums[g][m] = x;
        }
        lds_barrier();
#pragma unroll 8
        for (int g = half; g < RS_BATCH / 4; g += 2) acc = gxor(acc, sums[g][(mt[g][j][qw] >> qs) & 15]);
        lds_barrier();
    }
    // combine the two halves of each output bit
    if (half) other[b] = acc;
    __syncthreads();
    if (!half) partial[(uint64_t)blockIdx.x * 128 + b] = gxor(acc, other[b]);
}

// out[b] = XOR over blocks k of partial[k][b]: one block per output b, its
// lanes striding over the partial sums.
__global__ __launch_bounds__(BLOCK) void k_rs_reduce(const uint4 *__restrict__ partial, int n_blocks,
                                                     uint4 *__restrict__ out) {
    __shared__ uint4 red[BLOCK / 64];
    const int b = blockIdx.x;
    uint4 acc = make_uint4(0, 0, 0, 0);
    for (int k = threadIdx.x; k < n_blocks; k += BLOCK) acc = gxor(acc, partial[(uint64_t)k * 128 + b]);
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        acc.x ^= __shfl_xor(acc.x, m);
        acc.y ^= __shfl_xor(acc.y, m);
        acc.z ^= __shfl_xor(acc.z, m);
        acc.w ^= __shfl_xor(acc.w, m);
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < BLOCK / 64; w++) acc = gxor(acc, red[w]);
        out[b] = acc;
    }
}

int rs_fold_1b_rows_blocks(uint64_t n) {
    const uint64_t batches = (n + RS_BATCH - 1) / RS_BATCH;
    return (int)(batches < 2048 ? batches : 2048);
}

hipError_t launch_rs_fold_1b_rows(const uint4 *mat, const uint4 *vec, uint64_t n, uint4 *partial, uint4 *out,
                                  hipStream_t s) {
    const int blocks = rs_fold_1b_rows_blocks(n);
    k_rs_fold_1b_rows<<<blocks, BLOCK, 0, s>>>(mat, vec, n, partial);
    k_rs_reduce<<<128, BLOCK, 0, s>>>(partial, blocks, out);
    return hipGetLastError();
}

}  // namespace b64
