// block_decode.h -- per-block decoders of the four formats, one lane per
// block, pixels as 16 little-endian RGBA8 words in row-major order.
//
// Semantics follow the decoders the reference's quality harness uses:
// BC1 from Src/decoder.cpp:819-871 (5/6-bit expansion by bit replication,
// thirds by truncating integer division, 3-colour mode when c0 <= c1 with
// transparent black), ETC1 from Src/decoder.cpp:388-678 (individual and
// differential modes, eight intensity tables, MSB/LSB selector planes in
// column-major pixel order, flip bit), BC3 alpha from :874-925 (8 entries at
// sevenths when a0 > a1, else 6 at fifths plus 0 and 255, truncating) with
// its colour block always in 4-colour mode (:945-956), and EAC alpha from
// :679-713 with the tables of :25-42.  ETC1 blocks whose differential sums
// leave 0..31 are ETC2 T/H/planar blocks, which the encoder never produces;
// they are decoded here with 5-bit wrap-around and are outside the parity
// claim.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace goofy_dev {

__device__ __forceinline__ uint32_t pack_rgba(uint32_t r, uint32_t g, uint32_t b, uint32_t a)
{
    return r | (g << 8) | (b << 16) | (a << 24);
}

// 16 decoded pixels (row-major) of a BC1 block; FOUR: the colour block of
// BC3, always in 4-colour mode.
template <bool FOUR = false>
__device__ __forceinline__ void decode_bc1(uint32_t w0, uint32_t w1, uint32_t (&px)[16])
{
    const uint32_t c0 = w0 & 0xFFFFu, c1 = w0 >> 16;
    const uint32_t r0 = (c0 >> 11) & 31u, g0 = (c0 >> 5) & 63u, b0 = c0 & 31u;
    const uint32_t r1 = (c1 >> 11) & 31u, g1 = (c1 >> 5) & 63u, b1 = c1 & 31u;
    const uint32_t R0 = (r0 << 3) | (r0 >> 2), G0 = (g0 << 2) | (g0 >> 4), B0 = (b0 << 3) | (b0 >> 2);
    const uint32_t R1 = (r1 << 3) | (r1 >> 2), G1 = (g1 << 2) | (g1 >> 4), B1 = (b1 << 3) | (b1 >> 2);
    uint32_t pal[4];
    pal[0] = pack_rgba(R0, G0, B0, 255u);
    pal[1] = pack_rgba(R1, G1, B1, 255u);
    if (FOUR || c0 > c1) {
        pal[2] = pack_rgba((2u * R0 + R1) / 3u, (2u * G0 + G1) / 3u, (2u * B0 + B1) / 3u, 255u);
        pal[3] = pack_rgba((R0 + 2u * R1) / 3u, (G0 + 2u * G1) / 3u, (B0 + 2u * B1) / 3u, 255u);
    } else {
        pal[2] = pack_rgba((R0 + R1) / 2u, (G0 + G1) / 2u, (B0 + B1) / 2u, 255u);
        pal[3] = 0u;
    }
#pragma unroll
    for (int p = 0; p < 16; ++p) {
        const uint32_t k = (w1 >> (2 * p)) & 3u;
        const uint32_t lo = (k & 1u) ? pal[1] : pal[0];
        const uint32_t hi = (k & 1u) ? pal[3] : pal[2];
        px[p] = (k & 2u) ? hi : lo;
    }
}

__device__ __forceinline__ uint32_t clamp_add(uint32_t c, int32_t d)
{
    const int32_t v = (int32_t)c + d;
    return (uint32_t)(v < 0 ? 0 : (v > 255 ? 255 : v));
}

// 16 decoded pixels (row-major) of an ETC1 block; w0/w1 are the block's
// first/second 4 bytes read as little-endian words.
__device__ __forceinline__ void decode_etc1(uint32_t w0, uint32_t w1, uint32_t (&px)[16])
{
    const uint32_t hi = __builtin_bswap32(w0), lo = __builtin_bswap32(w1);  // big-endian words
    const bool diff = (hi >> 1) & 1u, flip = hi & 1u;
    uint32_t base[2][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const uint32_t field = (hi >> (24 - 8 * c)) & 0xFFu;
        if (diff) {
            const uint32_t a5 = field >> 3;
            const int32_t d = (int32_t)(field & 7u) - ((field & 4u) ? 8 : 0);
            const uint32_t b5 = (uint32_t)((int32_t)a5 + d) & 31u;
            base[0][c] = (a5 << 3) | (a5 >> 2);
            base[1][c] = (b5 << 3) | (b5 >> 2);
        } else {
            const uint32_t a4 = field >> 4, b4 = field & 15u;
            base[0][c] = (a4 << 4) | a4;
            base[1][c] = (b4 << 4) | b4;
        }
    }
    const uint32_t cw[2] = {(hi >> 5) & 7u, (hi >> 2) & 7u};
    uint32_t pal[2][4];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        // ETC1 intensity modifiers {small, large} per table index, one byte each:
        // {2,8},{5,17},{9,29},{13,42},{18,60},{24,80},{33,106},{47,183}
        const uint32_t a = (uint32_t)((0x2F2118120D090502ull >> (8 * cw[s])) & 0xFFu);
        const uint32_t b = (uint32_t)((0xB76A503C2A1D1108ull >> (8 * cw[s])) & 0xFFu);
        const int32_t mods[4] = {(int32_t)a, (int32_t)b, -(int32_t)a, -(int32_t)b};  // selector 0..3
#pragma unroll
        for (int k = 0; k < 4; ++k)
            pal[s][k] = pack_rgba(clamp_add(base[s][0], mods[k]), clamp_add(base[s][1], mods[k]),
                                  clamp_add(base[s][2], mods[k]), 255u);
    }
#pragma unroll
    for (int y = 0; y < 4; ++y)
#pragma unroll
        for (int x = 0; x < 4; ++x) {
            const int k = 4 * x + y;
            const uint32_t sel = (((lo >> (16 + k)) & 1u) << 1) | ((lo >> k) & 1u);
            const bool second = flip ? (y >= 2) : (x >= 2);
            const uint32_t* P = second ? pal[1] : pal[0];
            const uint32_t l = (sel & 1u) ? P[1] : P[0];
            const uint32_t h = (sel & 1u) ? P[3] : P[2];
            px[4 * y + x] = (sel & 2u) ? h : l;
        }
}

// Alpha of four pixels from an 8-entry palette held as bytes of {hi:lo}:
// v_perm_b32 with the four 3-bit codes as selector bytes.
__device__ __forceinline__ uint32_t lookup4(uint32_t pal_hi, uint32_t pal_lo, uint32_t codes)
{
    return __builtin_amdgcn_perm(pal_hi, pal_lo, codes);
}

// Replace byte 3 (alpha) of the four pixels of row r with the bytes of a4.
__device__ __forceinline__ void put_alpha(uint32_t (&px)[16], int r, uint32_t a4)
{
#pragma unroll
    for (int c = 0; c < 4; ++c) px[4 * r + c] = __builtin_amdgcn_perm(a4, px[4 * r + c], 0x04020100u + (c << 24));
}

// BC3 alpha block (words w0, w1) into byte 3 of px.
__device__ __forceinline__ void decode_bc3_alpha(uint32_t w0, uint32_t w1, uint32_t (&px)[16])
{
    const uint32_t a0 = w0 & 0xFFu, a1 = (w0 >> 8) & 0xFFu;
    uint32_t v[8];
    v[0] = a0;
    v[1] = a1;
    if (a0 > a1) {
#pragma unroll
        for (int i = 1; i < 7; ++i) v[1 + i] = ((7u - i) * a0 + i * a1) / 7u;
    } else {
#pragma unroll
        for (int i = 1; i < 5; ++i) v[1 + i] = ((5u - i) * a0 + i * a1) / 5u;
        v[6] = 0u;
        v[7] = 255u;
    }
    const uint32_t lo = v[0] | (v[1] << 8) | (v[2] << 16) | (v[3] << 24);
    const uint32_t hi = v[4] | (v[5] << 8) | (v[6] << 16) | (v[7] << 24);
    const uint64_t bits = (uint64_t)(w0 >> 16) | ((uint64_t)w1 << 16);  // pixel p at bits 3p
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint32_t g = (uint32_t)(bits >> (12 * r));
        const uint32_t codes = (g & 7u) | (((g >> 3) & 7u) << 8) | (((g >> 6) & 7u) << 16) | (((g >> 9) & 7u) << 24);
        put_alpha(px, r, lookup4(hi, lo, codes));
    }
}

// ETC2 alpha (EAC) modifier magnitudes: table t holds -m0..-m3 (selectors
// 0..3) and m0-1..m3-1 (selectors 4..7), m_k in nibble k of the 16-bit
// field t (the tables of Src/decoder.cpp:25-42 all have that shape).
__device__ __forceinline__ uint32_t eac_mags(uint32_t t)
{
    const uint64_t k[4] = {0xD642D852DA73F963ull, 0xB853B874B973C863ull, 0xA752A842A852A862ull,
                           0x97539864A321A743ull};
    return (uint32_t)(k[t >> 2] >> (16u * (t & 3u))) & 0xFFFFu;
}

// EAC alpha block (words w0, w1) into byte 3 of px.
__device__ __forceinline__ void decode_eac_alpha(uint32_t w0, uint32_t w1, uint32_t (&px)[16])
{
    const int32_t base = (int32_t)(w0 & 0xFFu), mult = (int32_t)((w0 >> 12) & 15u);
    const uint32_t mags = eac_mags((w0 >> 8) & 15u);
    uint32_t lo = 0, hi = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int32_t m = (int32_t)((mags >> (4 * k)) & 15u);
        lo |= (uint32_t)min(max(base - m * mult, 0), 255) << (8 * k);
        hi |= (uint32_t)min(max(base + (m - 1) * mult, 0), 255) << (8 * k);
    }
    // 48 selector bits, big-endian from byte 2: pixel (x, y) at 45 - 3(4x + y)
    const uint64_t bits = ((uint64_t)__builtin_bswap32(w0 & 0xFFFF0000u) << 32 | __builtin_bswap32(w1)) &
                          0xFFFFFFFFFFFFull;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        uint32_t codes = 0;
#pragma unroll
        for (int c = 0; c < 4; ++c) codes |= (uint32_t)((bits >> (45 - 3 * (4 * c + r))) & 7u) << (8 * c);
        put_alpha(px, r, lookup4(hi, lo, codes));
    }
}

template <int FMT>
__device__ __forceinline__ void decode_block(uint4 b, uint32_t (&px)[16])
{
    if constexpr (FMT == 0) decode_bc1(b.x, b.y, px);
    else if constexpr (FMT == 1) decode_etc1(b.x, b.y, px);
    else if constexpr (FMT == 2) {
        decode_bc1<true>(b.z, b.w, px);
        decode_bc3_alpha(b.x, b.y, px);
    } else {
        decode_etc1(b.z, b.w, px);
        decode_eac_alpha(b.x, b.y, px);
    }
}

}  // namespace goofy_dev
