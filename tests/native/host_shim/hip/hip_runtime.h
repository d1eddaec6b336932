// Host stand-in for <hip/hip_runtime.h>, used only to build the device block
// codec (goofy_amd/csrc/block_codec.h) as ordinary host C++ for a CPU test.
// It supplies what that header uses: the qualifiers, the uint2/uint4 vector
// types, min/max, and the two gfx950 byte instructions written from their
// documented semantics.
#pragma once

#include <stdint.h>

#include <algorithm>

#define __device__
#define __host__
#define __forceinline__ inline __attribute__((always_inline))

struct uint2 {
    uint32_t x, y;
};
struct uint4 {
    uint32_t x, y, z, w;
};
inline uint2 make_uint2(uint32_t x, uint32_t y) { return uint2{x, y}; }
inline uint4 make_uint4(uint32_t x, uint32_t y, uint32_t z, uint32_t w) { return uint4{x, y, z, w}; }
using std::max;
using std::min;

// V_PERM_B32: byte i of the result is byte sel_i of the 8-byte value
// {hi:lo} (0..3 from lo, 4..7 from hi); selectors 8..11 replicate the sign
// bit of byte 1, 3, 5, 7; 12 gives 0x00 and 13 and above 0xFF.
inline uint32_t goofy_host_perm(uint32_t hi, uint32_t lo, uint32_t sel)
{
    const uint64_t v = ((uint64_t)hi << 32) | lo;
    uint32_t out = 0;
    for (int i = 0; i < 4; ++i) {
        const uint32_t s = (sel >> (8 * i)) & 0xFFu;
        uint32_t b;
        if (s < 8) b = (uint32_t)(v >> (8 * s)) & 0xFFu;
        else if (s < 12) b = ((v >> (8 * (2 * (s - 8) + 1) + 7)) & 1u) ? 0xFFu : 0u;
        else b = s == 12 ? 0u : 0xFFu;
        out |= b << (8 * i);
    }
    return out;
}

// V_LERP_U8: per byte, (a + b + (bit 0 of c's byte)) >> 1.
inline uint32_t goofy_host_lerp(uint32_t a, uint32_t b, uint32_t c)
{
    uint32_t out = 0;
    for (int i = 0; i < 4; ++i) {
        const uint32_t s = ((a >> (8 * i)) & 0xFFu) + ((b >> (8 * i)) & 0xFFu) + ((c >> (8 * i)) & 1u);
        out |= ((s >> 1) & 0xFFu) << (8 * i);
    }
    return out;
}

// V_BITOP3_B32: bit i of the result is bit (a_i << 2 | b_i << 1 | c_i) of imm.
inline uint32_t goofy_host_bitop3(uint32_t a, uint32_t b, uint32_t c, uint32_t imm)
{
    uint32_t out = 0;
    for (int i = 0; i < 32; ++i) {
        const uint32_t k = (((a >> i) & 1u) << 2) | (((b >> i) & 1u) << 1) | ((c >> i) & 1u);
        out |= ((imm >> k) & 1u) << i;
    }
    return out;
}

// V_ALIGNBIT_B32: bits s..s+31 of the 64-bit value a:b (s taken mod 32).
inline uint32_t goofy_host_alignbit(uint32_t a, uint32_t b, uint32_t s)
{
    return (uint32_t)((((uint64_t)a << 32) | b) >> (s & 31u));
}

#define __builtin_amdgcn_perm goofy_host_perm
#define __builtin_amdgcn_alignbit goofy_host_alignbit
#define __builtin_amdgcn_lerp goofy_host_lerp
#define __builtin_amdgcn_bitop3_b32 goofy_host_bitop3
