// Three-input bitwise operations of gfx950 (v_bitop3_b32).  The immediate is
// the truth table f(0xF0, 0xCC, 0xAA) of the three operands in order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace b64 {

__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}
// m ? a : b, bit by bit
__device__ __forceinline__ uint32_t bsel(uint32_t a, uint32_t b, uint32_t m) {
    return __builtin_amdgcn_bitop3_b32(a, b, m, 0xE4);
}
// majority of three (SHA-256 Maj)
__device__ __forceinline__ uint32_t maj3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0xE8);
}

}  // namespace b64
