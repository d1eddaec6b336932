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
// :679-713 with the tables of :25-42.  A differential ETC block whose red,
// green or blue sum leaves 0..31 is an ETC2 T, H or planar block (checked in
// that order, Src/decoder.cpp:716-773; modes at :270-386); both ETC formats
// decode all five modes, as the harness does.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace goofy_dev {

// The decoders below hand out one pixel row at a time: emit(y, row) gets the
// four RGBA8 words of row y (y = 0..3, in order), so a caller that consumes
// rows as they come (a store, a comparison) keeps 4 pixels live, not 16.

// Replace byte 3 (alpha) of a row's four pixels with the bytes of a4.
__device__ __forceinline__ void put_alpha(uint32_t (&row)[4], uint32_t a4)
{
#pragma unroll
    for (int c = 0; c < 4; ++c) row[c] = __builtin_amdgcn_perm(a4, row[c], 0x04020100u + (c << 24));
}

// BC1 block rows; FOUR: the colour block of BC3, always in 4-colour mode,
// its alpha bytes from alpha_row(y).  Per channel the 4 palette bytes sit in
// one word (byte k = entry k); a row's pixels come from one v_perm_b32 per
// channel whose selector bytes are the pixels' 2-bit indices, and six more
// interleave the channel rows and the alpha row into RGBA words.
template <bool FOUR, class Alpha, class Emit>
__device__ __forceinline__ void bc1_rows(uint32_t w0, uint32_t w1, Alpha&& alpha_row, Emit&& emit)
{
    const uint32_t c0 = w0 & 0xFFFFu, c1 = w0 >> 16;
    const uint32_t r0 = (c0 >> 11) & 31u, g0 = (c0 >> 5) & 63u, b0 = c0 & 31u;
    const uint32_t r1 = (c1 >> 11) & 31u, g1 = (c1 >> 5) & 63u, b1 = c1 & 31u;
    const uint32_t e0[3] = {(r0 << 3) | (r0 >> 2), (g0 << 2) | (g0 >> 4), (b0 << 3) | (b0 >> 2)};
    const uint32_t e1[3] = {(r1 << 3) | (r1 >> 2), (g1 << 2) | (g1 >> 4), (b1 << 3) | (b1 >> 2)};
    const bool four = FOUR || c0 > c1;
    uint32_t plane[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const uint32_t e2 = four ? (2u * e0[c] + e1[c]) / 3u : (e0[c] + e1[c]) / 2u;
        const uint32_t e3 = four ? (e0[c] + 2u * e1[c]) / 3u : 0u;
        plane[c] = e0[c] | (e1[c] << 8) | (e2 << 16) | (e3 << 24);
    }
    const uint32_t alpha = four ? 0xFFFFFFFFu : 0x00FFFFFFu;  // 3-colour mode: entry 3 is transparent black
#pragma unroll
    for (int y = 0; y < 4; ++y) {
        const uint32_t x = (w1 >> (8 * y)) & 0xFFu;  // pixel c's index at bits 2c
        const uint32_t t = (x & 0x0Fu) | ((x & 0xF0u) << 12);
        const uint32_t sel = (t & 0x00030003u) | ((t & 0x000C000Cu) << 6);  // pixel c's index in byte c
        const uint32_t r = __builtin_amdgcn_perm(0u, plane[0], sel);
        const uint32_t g = __builtin_amdgcn_perm(0u, plane[1], sel);
        const uint32_t bl = __builtin_amdgcn_perm(0u, plane[2], sel);
        const uint32_t rg01 = __builtin_amdgcn_perm(g, r, 0x05010400u);   // R0 G0 R1 G1
        const uint32_t rg23 = __builtin_amdgcn_perm(g, r, 0x07030602u);   // R2 G2 R3 G3
        // alpha: the BC3 alpha block's row, or BC1's own (3-colour mode)
        const uint32_t a = FOUR ? alpha_row(y) : __builtin_amdgcn_perm(0u, alpha, sel);
        const uint32_t ba01 = __builtin_amdgcn_perm(a, bl, 0x05010400u);  // B0 A0 B1 A1
        const uint32_t ba23 = __builtin_amdgcn_perm(a, bl, 0x07030602u);  // B2 A2 B3 A3
        uint32_t row[4] = {__builtin_amdgcn_perm(ba01, rg01, 0x05040100u),
                           __builtin_amdgcn_perm(ba01, rg01, 0x07060302u),
                           __builtin_amdgcn_perm(ba23, rg23, 0x05040100u),
                           __builtin_amdgcn_perm(ba23, rg23, 0x07060302u)};
        emit(y, row);
    }
}

// Byte clamp of a small signed sum.
__device__ __forceinline__ uint32_t clamp_u8(int32_t v) { return (uint32_t)min(max(v, 0), 255); }

// ETC2 T mode (tmode) or H mode: the palette word of each channel's four
// paint colours, entry = 2 * MSB + LSB of a pixel's selector bits.  hi = the
// block's first four bytes as a big-endian word (block bits 63..32).
//   T: colour A red = bits 28-27 and 25-24 of hi, green 23-20, blue 19-16;
//      colour B red 15-12, green 11-8, blue 7-4; distance index = bits 3-2
//      then bit 0.  Paint colours A, B + d, B, B - d.
//   H: colour A red = 30-27, green 26-24 then 20, blue 19 then 17-15;
//      colour B red 14-11, green 10-7, blue 6-3; distance index = bit 2, bit
//      0, then 1 if A >= B as 12-bit RGB444 numbers.  Paint colours A + d,
//      A - d, B + d, B - d.
// Colours are 4 bits per channel widened by replication (x 17); distances
// 3, 6, 11, 16, 23, 32, 41, 64.
__device__ __forceinline__ void etc2_th_planes(uint32_t hi, bool tmode, uint32_t (&plane)[3])
{
    uint32_t a4[3], b4[3], di;
    if (tmode) {
        a4[0] = (((hi >> 27) & 3u) << 2) | ((hi >> 24) & 3u);
        a4[1] = (hi >> 20) & 15u;
        a4[2] = (hi >> 16) & 15u;
        b4[0] = (hi >> 12) & 15u;
        b4[1] = (hi >> 8) & 15u;
        b4[2] = (hi >> 4) & 15u;
        di = (((hi >> 2) & 3u) << 1) | (hi & 1u);
    } else {
        a4[0] = (hi >> 27) & 15u;
        a4[1] = (((hi >> 24) & 7u) << 1) | ((hi >> 20) & 1u);
        a4[2] = (((hi >> 19) & 1u) << 3) | ((hi >> 15) & 7u);
        b4[0] = (hi >> 11) & 15u;
        b4[1] = (hi >> 7) & 15u;
        b4[2] = (hi >> 3) & 15u;
        const uint32_t va = (a4[0] << 8) | (a4[1] << 4) | a4[2], vb = (b4[0] << 8) | (b4[1] << 4) | b4[2];
        di = (((hi >> 2) & 1u) << 2) | ((hi & 1u) << 1) | (va >= vb ? 1u : 0u);
    }
    const int32_t d = (int32_t)((0x40292017100B0603ull >> (8 * di)) & 0xFFu);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int32_t a = (int32_t)(a4[c] * 17u), b = (int32_t)(b4[c] * 17u);
        plane[c] = tmode ? (uint32_t)a | (clamp_u8(b + d) << 8) | ((uint32_t)b << 16) | (clamp_u8(b - d) << 24)
                         : clamp_u8(a + d) | (clamp_u8(a - d) << 8) | (clamp_u8(b + d) << 16) | (clamp_u8(b - d) << 24);
    }
}

// ETC2 planar mode: three RGB676 colours O, H, V (hi, lo = the block's two
// big-endian words); pixel (x, y) = clamp((x (H - O) + y (V - O) + 4 O + 2) >> 2),
// alpha 255.  O: red = bits 30-25 of hi, green 24 then 22-17, blue 16 then
// 12-11 then 9-7; H: red 6-2 then 0, green = bits 31-25 of lo, blue 24-19;
// V: red 18-13, green 12-6, blue 5-0.  6- and 7-bit values are widened by
// replicating their top bits.  Writes the four pixels of row y.
__device__ __forceinline__ void etc2_planar_row(uint32_t hi, uint32_t lo, int y, uint32_t (&row)[4])
{
    const uint32_t o6[3] = {(hi >> 25) & 63u, (((hi >> 24) & 1u) << 6) | ((hi >> 17) & 63u),
                            (((hi >> 16) & 1u) << 5) | (((hi >> 11) & 3u) << 3) | ((hi >> 7) & 7u)};
    const uint32_t h6[3] = {(((hi >> 2) & 31u) << 1) | (hi & 1u), lo >> 25, (lo >> 19) & 63u};
    const uint32_t v6[3] = {(lo >> 13) & 63u, (lo >> 6) & 127u, lo & 63u};
#pragma unroll
    for (int x = 0; x < 4; ++x) row[x] = 0xFF000000u;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int up = c == 1 ? 1 : 2, down = c == 1 ? 6 : 4;  // 7-bit green, 6-bit red and blue
        const int32_t o = (int32_t)((o6[c] << up) | (o6[c] >> down));
        const int32_t dh = (int32_t)((h6[c] << up) | (h6[c] >> down)) - o;
        const int32_t dv = (int32_t)((v6[c] << up) | (v6[c] >> down)) - o;
        int32_t t = 4 * o + 2 + y * dv;
        // clamp((t >> 2), 0, 255) as clamp(t, 0, 1023) >> 2 with a logical
        // shift: the shift-then-clamp form compiles to v_ashr_pk_u8_i32,
        // whose upper result bits the compiler assumes zero and gfx950 does
        // not clear
#pragma unroll
        for (int x = 0; x < 4; ++x, t += dh) row[x] |= ((uint32_t)min(max(t, 0), 1023) >> 2) << (8 * c);
    }
}

// ETC1 / ETC2 colour block rows; w0/w1 are the block's first/second 4 bytes
// read as little-endian words.
//
// Per channel, the 8 clamped palette bytes of both sub-blocks sit in two
// words (sub-block 0 in the low word, entry = selector value).  A row's four
// pixels then come from one v_perm_b32 per channel whose selector bytes are
// the pixels' 2-bit selectors plus 4 for sub-block 1, and four more
// v_perm_b32 interleave the channel rows into RGBA words (alpha 255 from
// selector byte 0x0d).  T and H blocks use the same lookup with their four
// paint colours in both words; planar blocks replace each row's pixels.
template <bool ALPHA, class Alpha, class Emit>
__device__ __forceinline__ void etc_rows(uint32_t w0, uint32_t w1, Alpha&& alpha_row, Emit&& emit)
{
    const uint32_t hi = __builtin_bswap32(w0), lo = __builtin_bswap32(w1);  // big-endian words
    const bool diff = (hi >> 1) & 1u, flip = hi & 1u;
    uint32_t base[2][3];
    uint32_t overflow = 0;  // bit c: channel c's differential sum leaves 0..31
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const uint32_t field = (hi >> (24 - 8 * c)) & 0xFFu;
        if (diff) {
            const uint32_t a5 = field >> 3;
            const int32_t d = (int32_t)(field & 7u) - ((field & 4u) ? 8 : 0);
            const uint32_t sum = (uint32_t)((int32_t)a5 + d);
            overflow |= (sum > 31u ? 1u : 0u) << c;
            const uint32_t b5 = sum & 31u;
            base[0][c] = (a5 << 3) | (a5 >> 2);
            base[1][c] = (b5 << 3) | (b5 >> 2);
        } else {
            const uint32_t a4 = field >> 4, b4 = field & 15u;
            base[0][c] = (a4 << 4) | a4;
            base[1][c] = (b4 << 4) | b4;
        }
    }
    // ETC2 planar mode (blue overflow alone): rows from O, H and V, on a path
    // of their own
    if (__builtin_expect(overflow == 4u, 0)) {
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            uint32_t row[4];
            etc2_planar_row(hi, lo, y, row);
            if constexpr (ALPHA) put_alpha(row, alpha_row(y));
            emit(y, row);
        }
        return;
    }
    // plane[c][s]: byte k = clamp(base[s][c] + modifier k), k = selector value:
    // [v + a, v + b, v - a, v - b] for the table's {small a, large b}.  In
    // 16-bit lanes [a, b]: a saturating-at-255 add (v_pk_add_u16 then
    // v_pk_min_u16), a saturating-at-0 subtract (v_pk_sub_u16 with clamp) and
    // one v_perm_b32 of the four low bytes.
    typedef unsigned short u16x2_t __attribute__((ext_vector_type(2)));
    uint32_t plane[3][2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const uint32_t cw = (hi >> (5 - 3 * s)) & 7u;
        // ETC1 intensity modifiers {small, large} per table index, one byte each:
        // {2,8},{5,17},{9,29},{13,42},{18,60},{24,80},{33,106},{47,183}; byte cw
        // of each 8-byte table by v_perm_b32, a into byte 0 and b into byte 2
        const uint32_t a = __builtin_amdgcn_perm(0x2F211812u, 0x0D090502u, 0x0c0c0c00u | cw);
        const uint32_t b = __builtin_amdgcn_perm(0xB76A503Cu, 0x2A1D1108u, 0x0c000c0cu | (cw << 16));
        const u16x2_t ab = __builtin_bit_cast(u16x2_t, a | b), top = {255, 255};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const u16x2_t v = __builtin_bit_cast(u16x2_t, base[s][c] * 0x10001u);
            const uint32_t up = __builtin_bit_cast(uint32_t, __builtin_elementwise_min(v + ab, top));
            const uint32_t dn = __builtin_bit_cast(uint32_t, __builtin_elementwise_sub_sat(v, ab));
            plane[c][s] = __builtin_amdgcn_perm(dn, up, 0x06040200u);
        }
    }
    // ETC2: red overflow = T mode, else green = H mode (else blue = planar, below)
    if (__builtin_expect((overflow & 3u) != 0u, 0)) {
        uint32_t th[3];
        etc2_th_planes(hi, (overflow & 1u) != 0u, th);
#pragma unroll
        for (int c = 0; c < 3; ++c) plane[c][0] = plane[c][1] = th[c];
    }
    // Selector bits of pixel (x, y) at k = 4x + y: LSB plane bit k, MSB plane
    // bit 16 + k, so column x's LSB and MSB nibbles sit at 4x and 16 + 4x.
    // One byte per column: [LSB nibble | MSB nibble << 4] in byte x (a byte
    // gather, then the middle nibbles of each 16-bit lane swapped); row y's
    // selector of column x is then bits y and 4 + y of byte x.
    const uint32_t g = __builtin_amdgcn_perm(lo, lo, 0x03010200u);  // [L0 L1], [M0 M1], [L2 L3], [M2 M3]
    const uint32_t cols = (g & 0xF00FF00Fu) | ((g & 0x00F000F0u) << 4) | ((g >> 4) & 0x00F000F0u);
    // sub-block 1: rows 2..3 (flip) or columns 2..3
#pragma unroll
    for (int y = 0; y < 4; ++y) {
        const uint32_t t = cols >> y;
        uint32_t sel = (t & 0x01010101u) | ((t >> 3) & 0x02020202u);  // pixel x's selector in byte x
        sel |= flip ? (y >= 2 ? 0x04040404u : 0u) : 0x04040000u;
        const uint32_t r = __builtin_amdgcn_perm(plane[0][1], plane[0][0], sel);
        const uint32_t g = __builtin_amdgcn_perm(plane[1][1], plane[1][0], sel);
        const uint32_t bl = __builtin_amdgcn_perm(plane[2][1], plane[2][0], sel);
        const uint32_t rg01 = __builtin_amdgcn_perm(g, r, 0x05010400u);   // R0 G0 R1 G1
        const uint32_t rg23 = __builtin_amdgcn_perm(g, r, 0x07030602u);   // R2 G2 R3 G3
        if constexpr (ALPHA) {
            const uint32_t a = alpha_row(y);
            const uint32_t ba01 = __builtin_amdgcn_perm(a, bl, 0x05010400u);  // B0 A0 B1 A1
            const uint32_t ba23 = __builtin_amdgcn_perm(a, bl, 0x07030602u);  // B2 A2 B3 A3
            uint32_t row[4] = {__builtin_amdgcn_perm(ba01, rg01, 0x05040100u),
                               __builtin_amdgcn_perm(ba01, rg01, 0x07060302u),
                               __builtin_amdgcn_perm(ba23, rg23, 0x05040100u),
                               __builtin_amdgcn_perm(ba23, rg23, 0x07060302u)};
            emit(y, row);
        } else {
            // pixel x = R_x G_x from rg, B_x from bl, alpha 255 (selector 0x0d)
            uint32_t row[4] = {__builtin_amdgcn_perm(bl, rg01, 0x0d040100u),
                               __builtin_amdgcn_perm(bl, rg01, 0x0d050302u),
                               __builtin_amdgcn_perm(bl, rg23, 0x0d060100u),
                               __builtin_amdgcn_perm(bl, rg23, 0x0d070302u)};
            emit(y, row);
        }
    }
}


// An alpha palette of 8 levels held as the bytes of {hi:lo}, and the block's
// 48 code bits; row r's four alphas are one v_perm_b32 with the row's 3-bit
// codes as selector bytes.
// An alpha block's palette (level k in byte k of lo | hi << 32) and its 48
// selector bits in two words, s0 and s1, laid out per format (below) so that
// a row's four 3-bit selectors come out with a handful of 32-bit operations.
struct AlphaPalette {
    uint32_t lo, hi;
    uint32_t s0, s1;
};

// BC3 alpha block (words w0, w1): 8 levels at sevenths when a0 > a1, else 6
// at fifths plus 0 and 255; codes little-endian, pixel p at bits 3p of the
// 48 bits after a0, a1 (row r's four codes at 12r .. 12r + 11).
__device__ __forceinline__ AlphaPalette bc3_palette(uint32_t w0, uint32_t w1)
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
    AlphaPalette p;
    p.lo = v[0] | (v[1] << 8) | (v[2] << 16) | (v[3] << 24);
    p.hi = v[4] | (v[5] << 8) | (v[6] << 16) | (v[7] << 24);
    p.s0 = (w0 >> 16) | (w1 << 16);  // selector bits 0..31
    p.s1 = w1 >> 16;                 // bits 32..47
    return p;
}

__device__ __forceinline__ uint32_t bc3_alpha_row(const AlphaPalette& p, int r)
{
    // row r's 12 bits at 12r of the 48 (a funnel shift of s1:s0), then the
    // four 3-bit fields spread to bytes 0..3: fields 0, 1 stay at bits 0..5
    // and fields 2, 3 move to 16..21, then fields 1 and 3 move up 5 bits
    const uint32_t g = r < 3 ? __builtin_amdgcn_alignbit(p.s1, p.s0, 12u * (uint32_t)r) : p.s1 >> 4;
    const uint32_t t = (g & 0x3Fu) | ((g & 0xFC0u) << 10);
    const uint32_t codes = (t & 0x00070007u) | ((t & 0x00380038u) << 5);
    return __builtin_amdgcn_perm(p.hi, p.lo, codes);
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

// EAC alpha block (words w0, w1): base, multiplier and table, then 48
// big-endian selector bits, pixel (x, y) at 45 - 3(4x + y): column c's 12
// bits at 36 - 12c .. 47 - 12c, row 0 in the top three.
__device__ __forceinline__ AlphaPalette eac_palette(uint32_t w0, uint32_t w1)
{
    const int32_t base = (int32_t)(w0 & 0xFFu), mult = (int32_t)((w0 >> 12) & 15u);
    const uint32_t mags = eac_mags((w0 >> 8) & 15u);
    AlphaPalette p;
    p.lo = 0;
    p.hi = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int32_t m = (int32_t)((mags >> (4 * k)) & 15u);
        p.lo |= clamp_u8(base - m * mult) << (8 * k);
        p.hi |= clamp_u8(base + (m - 1) * mult) << (8 * k);
    }
    // s0 / s1: columns 0, 1 / 2, 3 in the two 16-bit lanes, each column's 12
    // bits as they stand (row r at bit 9 - 3r)
    const uint32_t hi48 = __builtin_bswap32(w0 & 0xFFFF0000u), lo48 = __builtin_bswap32(w1);  // bits 32..47, 0..31
    const uint32_t c0 = (hi48 >> 4) & 0xFFFu, c1 = __builtin_amdgcn_alignbit(hi48, lo48, 24u) & 0xFFFu;
    const uint32_t c2 = (lo48 >> 12) & 0xFFFu, c3 = lo48 & 0xFFFu;
    p.s0 = c0 | (c1 << 16);
    p.s1 = c2 | (c3 << 16);
    return p;
}

__device__ __forceinline__ uint32_t eac_alpha_row(const AlphaPalette& p, int r)
{
    const uint32_t sh = 9u - 3u * (uint32_t)r;
    const uint32_t x = (p.s0 >> sh) & 0x00070007u, y = (p.s1 >> sh) & 0x00070007u;
    const uint32_t codes = __builtin_amdgcn_perm(y, x, 0x06040200u);  // [c0, c1, c2, c3]
    return __builtin_amdgcn_perm(p.hi, p.lo, codes);
}

// Rows of a block of format FMT (b: words x, y for the 8-byte formats; alpha
// block in x, y and colour block in z, w for DXT5 / ETC2 RGBA8).
template <int FMT, class Emit>
__device__ __forceinline__ void decode_block_rows(uint4 b, Emit&& emit)
{
    const auto opaque = [](int) { return 0xFFFFFFFFu; };
    if constexpr (FMT == 0) {
        bc1_rows<false>(b.x, b.y, opaque, emit);
    } else if constexpr (FMT == 1) {
        etc_rows<false>(b.x, b.y, opaque, emit);
    } else if constexpr (FMT == 2) {
        const AlphaPalette p = bc3_palette(b.x, b.y);
        bc1_rows<true>(b.z, b.w, [&](int y) { return bc3_alpha_row(p, y); }, emit);
    } else {
        const AlphaPalette p = eac_palette(b.x, b.y);
        etc_rows<true>(b.z, b.w, [&](int y) { return eac_alpha_row(p, y); }, emit);
    }
}

// The 16 pixels of a block, row-major.
template <int FMT>
__device__ __forceinline__ void decode_block(uint4 b, uint32_t (&px)[16])
{
    decode_block_rows<FMT>(b, [&](int y, const uint32_t (&row)[4]) {
#pragma unroll
        for (int c = 0; c < 4; ++c) px[4 * y + c] = row[c];
    });
}

}  // namespace goofy_dev
