// Host-side helpers shared by the C ABI sources (api.hip, batch.hip, pcs.hip):
// stream casts, element packing and the status-code macros.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <vector>

#include "../../include/b64hip.h"
#include "host_util.hpp"

namespace b64 {
namespace abi {

using host::u128;

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }
inline u128 ld128(const uint64_t *p) { return ((u128)p[1] << 64) | p[0]; }
inline void st128(uint64_t *p, u128 v) {
    p[0] = (uint64_t)v;
    p[1] = (uint64_t)(v >> 64);
}
inline uint4 to_u4(u128 v) {
    return make_uint4((uint32_t)v, (uint32_t)(v >> 32), (uint32_t)(v >> 64), (uint32_t)(v >> 96));
}
inline uint4 to_u4(const uint64_t *p) { return to_u4(ld128(p)); }
inline u128 from_u4(uint4 v) {
    return ((u128)v.w << 96) | ((u128)v.z << 64) | ((u128)v.y << 32) | (u128)v.x;
}
inline int log2_ceil_u64(uint64_t x) {
    int r = 0;
    while ((1ull << r) < x) r++;
    return r;
}
inline int hip_status(hipError_t e) {
    return e == hipSuccess ? B64_OK : e == hipErrorOutOfMemory ? B64_ERR_OUT_OF_MEMORY : B64_ERR_HIP;
}

// Scratch buffers of the provers (folded oracles, their trees, reduction
// partials, gather lists) come from the device's default stream-ordered pool.
// Left at its default release threshold the pool hands freed memory back to the
// driver at every synchronisation, and each allocation then costs a full
// driver allocation again; keeping up to this many freed bytes cached makes the
// per-round allocations cheap.
constexpr uint64_t POOL_KEEP_BYTES = 8ull << 30;
inline void keep_pool_memory() {
    static std::mutex mu;
    static bool done[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return;
    std::lock_guard<std::mutex> lock(mu);
    if (done[dev]) return;
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
        uint64_t keep = POOL_KEEP_BYTES;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
    done[dev] = true;
}

// Small pinned host blocks (PINNED_SMALL bytes) for per-prover results that
// kernels or copies write while the host is not waiting: a process-wide free
// list, so a prover created per proof does not pay a pinned allocation.
constexpr size_t PINNED_SMALL = 256;
struct PinnedSmallPool {
    std::mutex mu;
    std::vector<void *> free_blocks;
};
inline PinnedSmallPool &pinned_small_pool() {
    static PinnedSmallPool *pool = new PinnedSmallPool();  // lives until the process ends
    return *pool;
}
inline void *pinned_small_get() {
    PinnedSmallPool &p = pinned_small_pool();
    {
        std::lock_guard<std::mutex> lock(p.mu);
        if (!p.free_blocks.empty()) {
            void *b = p.free_blocks.back();
            p.free_blocks.pop_back();
            return b;
        }
    }
    void *b = nullptr;
    return hipHostMalloc(&b, PINNED_SMALL, hipHostMallocDefault) == hipSuccess ? b : nullptr;
}
inline void pinned_small_put(void *b) {
    if (!b) return;
    PinnedSmallPool &p = pinned_small_pool();
    std::lock_guard<std::mutex> lock(p.mu);
    p.free_blocks.push_back(b);
}

}  // namespace abi
}  // namespace b64

#define HIP_TRY(expr)                                     \
    do {                                                  \
        const int _rc = b64::abi::hip_status((expr));     \
        if (_rc != B64_OK) return _rc;                    \
    } while (0)

#define B64_TRY(expr)                  \
    do {                               \
        const int _rc = (expr);        \
        if (_rc != B64_OK) return _rc; \
    } while (0)
