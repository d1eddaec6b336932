// block_codec.h -- the per-4x4-block DXT1 / ETC1S encoder for gfx950.
//
// One lane owns one block: its 16 RGBA8 pixels sit in 16 VGPRs (dword p =
// pixel 4*row+col, R in bits 0..7) and every step below is lane-local byte
// arithmetic.  This replaces the 16-byte SSE2 lanes of goofySimdEncode
// (GoofyTC/goofy_tc.h:1069-1494), which encode 4 blocks per call and spend
// most of their instructions on transposes (transposeAs4x4 :263-276,
// deinterleaveRGB :278-299, transposeAs4x4x4 :327-358); here there is no
// transpose at all: v_perm_b32 gathers bytes, v_lerp_u8 is the per-byte
// rounding average (_mm_avg_epu8), packed 16-bit min/max (three-input ones
// where the lanes hold bytes) give the bounding box, and per-byte comparisons
// come out of the same rounding average.
//
// Integer semantics (all values u8, avg(a,b) = (a+b+1)>>1):
//   Y(r,g,b) = avg(avg(r,b),g)                       goofy_tc.h:1166,1205
//   range = max(Y(max)-Y(min), 8), mid = avg(Y(min),Y(max))   :1176-1179
//   qt = avg(avg(range,0),0) + avg(avg(avg(range,0),0),0)     :1184-1190
//   gez(p) = Y(p) >= mid ; near(p) = |Y(p)-mid| < qt          :1212-1224
//     (the header clamps |Y-mid| to 127 before a signed compare; qt <= 96,
//      so the clamp never changes the outcome)
//   q5(c) = avg(avg(avg(subsat(c,8),0),0),0) = c ? (c-1)>>3 : 0   :1309
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace goofy_dev {

// DXT5 (BC3) and ETC2 RGBA8 are 16-byte blocks: an alpha block, then the
// DXT1 / ETC1S block of the same pixels.
enum : int { kDXT1 = 0, kETC1 = 1, kDXT5 = 2, kETC2 = 3 };

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

// Per-byte rounding average of two packed words: v_lerp_u8 with a carry-in
// of one in every byte, i.e. (a+b+1)>>1 per byte.
__device__ __forceinline__ uint32_t avg4(uint32_t a, uint32_t b) { return __builtin_amdgcn_lerp(a, b, 0x01010101u); }
// a ^ b ^ c as one v_bitop3_b32 (truth table 0x96, the same for every operand order).
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c)
{
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96u);
}

// Bits of m where `sel` is set, bits of o elsewhere: one v_bitop3_b32
// (truth table indexed by sel*4 + m*2 + o).  Written as the builtin so that
// the compiler keeps the single select instead of re-deriving masks from
// the known-zero bits of shifted operands.
__device__ __forceinline__ uint32_t bsel(uint32_t sel, uint32_t m, uint32_t o)
{
    return __builtin_amdgcn_bitop3_b32(sel, m, o, 0xCAu);
}

// v_perm_b32: byte i of the result is byte sel[i] of {hi:lo} (0..3 = lo,
// 4..7 = hi, 0x0c = zero).
__device__ __forceinline__ uint32_t perm(uint32_t hi, uint32_t lo, uint32_t sel)
{
    return __builtin_amdgcn_perm(hi, lo, sel);
}

__device__ __forceinline__ uint32_t pk_min16(uint32_t a, uint32_t b)
{
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_bit_cast(u16x2, a), __builtin_bit_cast(u16x2, b)));
}
__device__ __forceinline__ uint32_t pk_max16(uint32_t a, uint32_t b)
{
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(u16x2, a), __builtin_bit_cast(u16x2, b)));
}

// Three-input min / max of 16-bit lanes that hold 0..255, as one
// v_pk_minimum3_f16 / v_pk_maximum3_f16: read as half-precision numbers such
// lanes are +0 and positive subnormals, whose order is that of the integers,
// and the kernels keep half-precision subnormals (denorm mode 16/64 = 3), so
// the selected lane comes out bit for bit.  One instruction takes two new
// pairs of lanes where v_pk_min_u16 takes one.
typedef _Float16 h16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t min3_b16(uint32_t a, uint32_t b, uint32_t c)
{
    const h16x2 r = __builtin_elementwise_minimum(
        __builtin_elementwise_minimum(__builtin_bit_cast(h16x2, a), __builtin_bit_cast(h16x2, b)),
        __builtin_bit_cast(h16x2, c));
    return __builtin_bit_cast(uint32_t, r);
}
__device__ __forceinline__ uint32_t max3_b16(uint32_t a, uint32_t b, uint32_t c)
{
    const h16x2 r = __builtin_elementwise_maximum(
        __builtin_elementwise_maximum(__builtin_bit_cast(h16x2, a), __builtin_bit_cast(h16x2, b)),
        __builtin_bit_cast(h16x2, c));
    return __builtin_bit_cast(uint32_t, r);
}

// Luma of the colour in byte lanes 0..2 of c (result in byte 0, other bytes
// are don't-care).
__device__ __forceinline__ uint32_t luma_b0(uint32_t c)
{
    const uint32_t rb = avg4(c, c >> 16);  // byte0 = avg(R,B)
    return avg4(rb, c >> 8) & 0xFFu;       // byte0 = avg(avg(R,B),G)
}

// Luma of four pixels, packed: byte i = Y(p_i).
__device__ __forceinline__ uint32_t luma4(uint32_t p0, uint32_t p1, uint32_t p2, uint32_t p3)
{
    // [R0 R1 G0 G1] and [B0 B1 G0 G1] -> avg -> [avg(R,B)0, avg(R,B)1, G0, G1]
    const uint32_t q01 = avg4(perm(p1, p0, 0x05010400u), perm(p1, p0, 0x05010602u));
    const uint32_t q23 = avg4(perm(p3, p2, 0x05010400u), perm(p3, p2, 0x05010602u));
    // [rb0 rb1 rb2 rb3] and [G0 G1 G2 G3]
    return avg4(perm(q23, q01, 0x05040100u), perm(q23, q01, 0x07060302u));
}

// Byte 0 of x copied into all four bytes (v_perm_b32 broadcast; avoids the
// quarter-rate v_mul_lo_u32 of x * 0x01010101).
__device__ __forceinline__ uint32_t rep4(uint32_t x) { return perm(0u, x, 0x00000000u); }

// 8-bit channel -> 5-bit, closed form of the header's three-halvings cascade.
__device__ __forceinline__ uint32_t q5(uint32_t c) { return (c - (c != 0u ? 1u : 0u)) >> 3; }

struct BlockStats {
    uint32_t range, mid, qt;
    uint32_t mnrb, mxrb;  // R, B of the bounding-box corners in 16-bit lanes
    uint32_t mng, mxg;    // G of the corners
};

// The rest of the statistics from the bounding-box corners: R, B in the
// 16-bit lanes of mnrb / mxrb, G in mng / mxg.  Both corners' lumas come
// out of one pass of byte averages (bytes 0 and 1 of y2), and the roundings
// (x + 1) >> 1 of mid and qt are byte averages too (range <= 255).
__device__ __forceinline__ BlockStats corner_stats(uint32_t mnrb, uint32_t mxrb, uint32_t mng, uint32_t mxg)
{
    BlockStats s;
    s.mnrb = mnrb;
    s.mxrb = mxrb;
    s.mng = mng;
    s.mxg = mxg;
    const uint32_t rb = perm(mxrb, mnrb, 0x06020400u);             // [R_mn R_mx B_mn B_mx]
    const uint32_t y2 = avg4(avg4(rb, rb >> 16), mng | (mxg << 8));  // [Y(mn) Y(mx) . .]
    const uint32_t d = ((y2 >> 8) & 0xFFu) - (y2 & 0xFFu);          // >= 0: luma is monotone per channel
    s.range = d > 8u ? d : 8u;
    s.mid = avg4(y2, y2 >> 8) & 0xFFu;
    const uint32_t half = avg4(s.range, 0u);
    const uint32_t quarter = avg4(half, 0u);
    const uint32_t eighth = avg4(quarter, 0u);
    s.qt = quarter + eighth;  // <= 96
    return s;
}

// R, B bounding box in the two 16-bit halves of p & 0x00ff00ff: the 16
// masked words through a chain of three-input min / max (8 + 8 instructions
// instead of 15 + 15).
__device__ __forceinline__ void rb_box(const uint32_t (&p)[16], uint32_t& mnrb, uint32_t& mxrb)
{
    uint32_t rb[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) rb[i] = p[i] & 0x00FF00FFu;
    mnrb = min3_b16(rb[0], rb[1], rb[2]);
    mxrb = max3_b16(rb[0], rb[1], rb[2]);
#pragma unroll
    for (int i = 3; i < 15; i += 2) {
        mnrb = min3_b16(mnrb, rb[i], rb[i + 1]);
        mxrb = max3_b16(mxrb, rb[i], rb[i + 1]);
    }
    mnrb = pk_min16(mnrb, rb[15]);
    mxrb = pk_max16(mxrb, rb[15]);
}

__device__ __forceinline__ BlockStats block_stats(const uint32_t (&p)[16])
{
    uint32_t mnrb, mxrb;
    rb_box(p, mnrb, mxrb);
    // G of two pixels packed into one word with v_perm ([G_a 0 G_b 0]),
    // then the three-input min / max over the eight words
    uint32_t g[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) g[i] = perm(p[2 * i + 1], p[2 * i], 0x0c050c01u);
    uint32_t mng = min3_b16(g[0], g[1], g[2]), mxg = max3_b16(g[0], g[1], g[2]);
    mng = min3_b16(mng, g[3], g[4]);
    mxg = max3_b16(mxg, g[3], g[4]);
    mng = min3_b16(mng, g[5], g[6]);
    mxg = max3_b16(mxg, g[5], g[6]);
    mng = pk_min16(mng, g[7]);
    mxg = pk_max16(mxg, g[7]);
    mng = pk_min16(mng, mng >> 16) & 0xFFu;
    mxg = pk_max16(mxg, mxg >> 16) & 0xFFu;
    return corner_stats(mnrb, mxrb, mng, mxg);
}

// The same for a block whose alpha range is needed too: v_pk_min_u16 /
// v_pk_max_u16 over the raw pixel words order each 16-bit lane by its high
// byte first, so the lanes' high bytes come out as the G range (byte 1) and
// the alpha range (byte 3) together, in the 30 instructions that the G
// pairs and a separate alpha pass would take 44.
__device__ __forceinline__ BlockStats block_stats_alpha(const uint32_t (&p)[16], uint32_t& amn, uint32_t& amx)
{
    uint32_t mnrb, mxrb;
    rb_box(p, mnrb, mxrb);
    uint32_t lo = p[0], hi = p[0];
#pragma unroll
    for (int i = 1; i < 16; ++i) {
        lo = pk_min16(lo, p[i]);
        hi = pk_max16(hi, p[i]);
    }
    amn = lo >> 24;
    amx = hi >> 24;
    return corner_stats(mnrb, mxrb, (lo >> 8) & 0xFFu, (hi >> 8) & 0xFFu);
}

// Per-byte comparison masks for four packed lumas y4 against the block's mid
// and threshold: bit 7 of each byte of *gez = (y >= mid), of *near =
// (|y - mid| < qt) = (lo <= y <= hi).  Uses avg(~a, b) >= 128  <=>  b >= a
// for bytes a, b, and for the upper bound the floor average:
// (~hi + y) >> 1 >= 128  <=>  y > hi, which also holds for hi = 255 (never).
// nhi4 = ~hi in all four bytes.
__device__ __forceinline__ void compare4(uint32_t y4, uint32_t nmid4, uint32_t nlo4, uint32_t nhi4, uint32_t* gez,
                                         uint32_t* near)
{
    *gez = avg4(nmid4, y4);
    *near = avg4(nlo4, y4) & ~__builtin_amdgcn_lerp(nhi4, y4, 0u);
}

// q5 of the two 16-bit lanes of x: subsat(c, 1) >> 3 per lane.
__device__ __forceinline__ uint32_t q5x2(uint32_t x)
{
    const u16x2 one = {1, 1}, three = {3, 3};
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_sub_sat(__builtin_bit_cast(u16x2, x), one) >> three);
}

// The DXT1 endpoint word c0 | c1 << 16: c0 = RGB565 of the box maximum with
// the green LSB forced to 1 (so c0 > c1, 4-colour mode), c1 = RGB565 of the
// minimum.  All six channels go through q5 in 16-bit lanes: [R, B] of the
// maximum, [R, B] of the minimum and [G max, G min]; two v_perm_b32 pair the
// reds and the blues of both endpoints.
__device__ __forceinline__ uint32_t dxt1_endpoints(const BlockStats& s)
{
    const uint32_t x = q5x2(s.mxrb), n = q5x2(s.mnrb), g = q5x2(s.mxg | (s.mng << 16));
    const uint32_t r = perm(n, x, 0x0c040c00u);  // [R max, R min] in the two lanes
    const uint32_t b = perm(n, x, 0x0c060c02u);  // [B max, B min]
    return (r << 11) | (g << 6) | b | 0x20u;
}

// The DXT1 index word from the per-row compare masks (bit 7 of byte c of
// gez[r] / near[r] for pixel (r, c)): pixel (r, c)'s 2-bit code, bit 0 =
// y < mid and bit 1 = near, at bits 8r + 2c.  Per row, bit selects put both
// flags in bits 6, 7 of each byte, the byte pairs' codes side by side in bits
// 6..9 and 22..25, and all four in bits 6..13.
__device__ __forceinline__ uint32_t dxt1_index_word(const uint32_t (&gez)[4], const uint32_t (&near)[4])
{
    uint32_t idx = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        // near in bit 7, ~gez in bit 6 (truth table 0xC5: sel ? near : ~gez)
        const uint32_t z = __builtin_amdgcn_bitop3_b32(0x80808080u, near[r], gez[r] >> 1, 0xC5u);
        const uint32_t y = bsel(0x00C000C0u, z, z >> 6);
        const uint32_t g = bsel(0x000003C0u, y, y >> 12);
        idx |= ((g >> 6) & 0xFFu) << (8 * r);
    }
    return idx;
}

// The ETC1S selector word (goofy_tc.h:1375-1395: word bit j <-> column
// {2,3,0,1}[j>>2], row j&3; the MSB plane N in bits 16..31, then inverted)
// from the same masks.  Each plane gathers row r's flags into bit 4 + r of
// byte c (row 3's stay in bit 7), then a bit select folds the column nibbles
// into bytes 1 (columns 0, 1) and 3 (columns 2, 3).
__device__ __forceinline__ uint32_t etc1_selector_word(const uint32_t (&gez)[4], const uint32_t (&near)[4])
{
    auto plane = [](const uint32_t (&f)[4]) {
        const uint32_t t = bsel(0x10101010u, f[0] >> 3, bsel(0x20202020u, f[1] >> 2, bsel(0x40404040u, f[2] >> 1, f[3])));
        return bsel(0x0F000F00u, t << 4, t);
    };
    return ~perm(plane(near), plane(gez), 0x05070103u);  // G: cols 2|3, 0|1; then N's
}

// DXT1 (BC1) block: returns {c0 | c1<<16, indices} as two little-endian words.
__device__ __forceinline__ uint2 encode_dxt1(const uint32_t (&p)[16], const BlockStats& s)
{
    const uint32_t lo = s.mid + 1u > s.qt ? s.mid + 1u - s.qt : 0u;          // y >= lo
    const uint32_t hi = s.mid + s.qt - 1u < 255u ? s.mid + s.qt - 1u : 255u;  // y <= hi
    const uint32_t nmid4 = rep4(~s.mid);
    const uint32_t nlo4 = rep4(~lo);
    const uint32_t nhi4 = rep4(~hi);

    uint32_t gez[4], near[4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
        compare4(luma4(p[4 * r], p[4 * r + 1], p[4 * r + 2], p[4 * r + 3]), nmid4, nlo4, nhi4, &gez[r], &near[r]);
    return make_uint2(dxt1_endpoints(s), dxt1_index_word(gez, near));
}

__device__ __forceinline__ uint2 encode_dxt1(const uint32_t (&p)[16]) { return encode_dxt1(p, block_stats(p)); }

// ETC1 base-colour bytes of the block, from its average colour: each
// channel of avgc moved by the luma correction (up: + corr, else - corr) with
// saturation to 0..255 (goofy_tc.h:1431-1449), then cut to 5 bits in place,
// q5(b) << 3 == subsat(b, 1) & 0xf8.  The correction is the signed
// clamp(mid - Y(avgc), -127, 127), so a channel becomes
// clamp(c + clamp(mid - Y, -127, 127), 0, 255); with the subtraction of one
// folded in, clamp(c + clamp(mid - Y, -127, 127) - 1, 0, 254) & 0xf8.  R and
// B run in the two 16-bit lanes of one word, G in another: one signed 16-bit
// add and a v_pk_max_i16 / v_pk_min_i16 clamp per word.
typedef short i16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t etc1_base_bytes(uint32_t avgc, uint32_t mid)
{
    const int32_t ya = (int32_t)luma_b0(avgc);
    const int32_t t = (int32_t)mid - ya - 1;
    const int32_t sd = t < -128 ? -128 : (t > 126 ? 126 : t);  // the correction minus one
    const uint32_t sd2 = perm(0u, (uint32_t)sd, 0x01000100u);   // in both lanes
    const i16x2 zero = {0, 0}, top = {254, 254};
    auto base = [&](uint32_t lanes) {  // lanes: two channels in 16-bit lanes, 0..255 each
        const i16x2 v = __builtin_bit_cast(i16x2, lanes) + __builtin_bit_cast(i16x2, sd2);
        const i16x2 c = __builtin_elementwise_min(__builtin_elementwise_max(v, zero), top);
        return __builtin_bit_cast(uint32_t, c) & 0x00F800F8u;
    };
    return base(avgc & 0x00FF00FFu) | ((base((avgc >> 8) & 0x00FF00FFu) & 0xFFu) << 8);
}

// ETC1 codeword-table index for the brightness range: the number of the
// breakpoints 21, 43, 73, 105, 151, 181, 253 of etc1BrighnessRangeTocontrolByte
// (GoofyTC/goofy_tc.h:1040-1056) that the range exceeds.  range <= 255, so
// the seven compares range >= t + 1 run as byte compares, four to a word
// (avg(~t', r) >= 128 <=> r >= t'), and a popcount of their flags counts them.
__device__ __forceinline__ uint32_t etc1_table(uint32_t range)
{
    const uint32_t r4 = rep4(range);
    const uint32_t lo = avg4(~0x6A4A2C16u, r4);  // t + 1 = 22, 44, 74, 106
    const uint32_t hi = avg4(~0x00FEB698u, r4);  // 152, 182, 254 (byte 3 unused)
    return (uint32_t)(__builtin_popcount(lo & 0x80808080u) + __builtin_popcount(hi & 0x00808080u));
}

// ETC1S block: differential mode, zero deltas, one table for both halves,
// flip = 1; returns the 8 bytes as two little-endian words.
__device__ __forceinline__ uint2 encode_etc1(const uint32_t (&p)[16], const BlockStats& s)
{
    const uint32_t lo = s.mid + 1u > s.qt ? s.mid + 1u - s.qt : 0u;
    const uint32_t hi = s.mid + s.qt - 1u < 255u ? s.mid + s.qt - 1u : 255u;
    const uint32_t nmid4 = rep4(~s.mid);
    const uint32_t nlo4 = rep4(~lo);
    const uint32_t nhi4 = rep4(~hi);

    uint32_t gez[4], near[4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
        compare4(luma4(p[4 * r], p[4 * r + 1], p[4 * r + 2], p[4 * r + 3]), nmid4, nlo4, nhi4, &gez[r], &near[r]);
    const uint32_t w1 = etc1_selector_word(gez, near);

    // Average colour: rows first, then columns (goofy_tc.h:1402-1414).
    uint32_t col[4];
#pragma unroll
    for (int x = 0; x < 4; ++x) col[x] = avg4(avg4(p[x], p[4 + x]), avg4(p[8 + x], p[12 + x]));
    const uint32_t avgc = avg4(avg4(col[0], col[1]), avg4(col[2], col[3]));
    // Shift its brightness onto mid (goofy_tc.h:1431-1449).
    uint32_t w0 = etc1_base_bytes(avgc, s.mid);
    const uint32_t cw = etc1_table(s.range);
    w0 |= ((cw << 5) | (cw << 2) | 3u) << 24;
    return make_uint2(w0, w1);
}

__device__ __forceinline__ uint2 encode_etc1(const uint32_t (&p)[16]) { return encode_etc1(p, block_stats(p)); }

// ---- both formats from one classification (the dual-format kernel) --------
// The same arithmetic as encode_dxt1 / encode_etc1 above, split into the part
// both formats share -- statistics and, per row r, the per-byte masks of
// (y >= mid) and (|y - mid| < qt) -- and the packing of each format.  The
// single-format encoders keep their own instruction order.
//
// What both formats need from a block.
struct BlockClass {
    BlockStats s;
    uint32_t gez[4], near[4];
};

__device__ __forceinline__ BlockClass classify(const uint32_t (&p)[16], const BlockStats& s)
{
    BlockClass c;
    c.s = s;
    const uint32_t lo = c.s.mid + 1u > c.s.qt ? c.s.mid + 1u - c.s.qt : 0u;          // y >= lo
    const uint32_t hi = c.s.mid + c.s.qt - 1u < 255u ? c.s.mid + c.s.qt - 1u : 255u;  // y <= hi
    const uint32_t nmid4 = rep4(~c.s.mid);
    const uint32_t nlo4 = rep4(~lo);
    const uint32_t nhi4 = rep4(~hi);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint32_t y4 = luma4(p[4 * r], p[4 * r + 1], p[4 * r + 2], p[4 * r + 3]);
        compare4(y4, nmid4, nlo4, nhi4, &c.gez[r], &c.near[r]);
    }
    return c;
}

__device__ __forceinline__ BlockClass classify(const uint32_t (&p)[16]) { return classify(p, block_stats(p)); }

// DXT1 words from a classified block (as encode_dxt1).
__device__ __forceinline__ uint2 pack_dxt1(const BlockClass& c)
{
    return make_uint2(dxt1_endpoints(c.s), dxt1_index_word(c.gez, c.near));
}

// ETC1S words from a classified block and its pixels (as encode_etc1).
__device__ __forceinline__ uint2 pack_etc1(const uint32_t (&p)[16], const BlockClass& c)
{
    const uint32_t w1 = etc1_selector_word(c.gez, c.near);

    // Average colour: rows first, then columns (goofy_tc.h:1402-1414).
    uint32_t col[4];
#pragma unroll
    for (int x = 0; x < 4; ++x) col[x] = avg4(avg4(p[x], p[4 + x]), avg4(p[8 + x], p[12 + x]));
    const uint32_t avgc = avg4(avg4(col[0], col[1]), avg4(col[2], col[3]));
    // Shift its brightness onto mid (goofy_tc.h:1431-1449).
    uint32_t w0 = etc1_base_bytes(avgc, c.s.mid);
    const uint32_t cw = etc1_table(c.s.range);
    w0 |= ((cw << 5) | (cw << 2) | 3u) << 24;
    return make_uint2(w0, w1);
}

template <int FMT>
__device__ __forceinline__ uint2 encode_block(const uint32_t (&p)[16])
{
    if constexpr (FMT == kDXT1) return encode_dxt1(p);
    else return encode_etc1(p);
}

// ---- alpha blocks (DXT5 / ETC2 RGBA8) ---------------------------------------
// The reference has no alpha encoder; its README names one that works on
// alpha directly instead of brightness as the next step.  This one does that
// with the same byte arithmetic as the colour blocks:
//   * min/max alpha of the block;
//   * the 8 palette levels in ascending order, exactly as a decoder forms
//     them (Src/decoder.cpp:25-42 EAC tables, :874-896 BC3 palette);
//   * per pixel, its rank among the levels = the number of midpoints
//     ceil((lv[k-1] + lv[k]) / 2) it reaches, so every pixel takes the level
//     nearest to it and a pixel exactly between two takes the higher one;
//   * the rank's selector, packed in the format's bit order.
// BC3: a0 = max, a1 = min (8-level mode when they differ), level k =
// min + floor(k*(max-min)/7).  EAC: table 9 ({-10,-8,-5,-2,1,4,7,9} sorted),
// multiplier m = max(1, ceil((max-min)/19)) so that 19m covers the range,
// base min + 10m (lowest level on the minimum) or, when that passes 255,
// max - 9m (highest level on the maximum).

struct AlphaRows {
    uint32_t w[4];  // word r: byte c = alpha of pixel (row r, column c)
    uint32_t mn, mx;
};

// Alpha rows of a block whose alpha range [amn, amx] is known.
__device__ __forceinline__ AlphaRows alpha_rows(const uint32_t (&p)[16], uint32_t amn, uint32_t amx)
{
    AlphaRows a;
    a.mn = amn;
    a.mx = amx;
#pragma unroll
    for (int r = 0; r < 4; ++r)  // [A0 A1 0 0] | [0 0 A2 A3]
        a.w[r] = perm(p[4 * r + 1], p[4 * r], 0x0c0c0703u) | perm(p[4 * r + 3], p[4 * r + 2], 0x07030c0cu);
    return a;
}

// From the inverted ascending levels n_k = 255 - lv[k], each replicated to
// all four bytes: ~thr_k in all four bytes, thr_k = ceil((lv[k-1] + lv[k]) /
// 2).  255 - ceil(y / 2) = floor((510 - y) / 2), the floor average of n_{k-1}
// and n_k, which is v_lerp_u8 with a zero rounding operand.
__device__ __forceinline__ void rep_midpoints(const uint32_t (&r)[8], uint32_t (&nthr4)[7])
{
#pragma unroll
    for (int k = 1; k < 8; ++k) nthr4[k - 1] = __builtin_amdgcn_lerp(r[k - 1], r[k], 0u);
}

// Threshold flags of four alphas: m[k] has bit 7 of byte c set when alpha c
// reaches thr_{k+1} (f_{k+1} below); the other bits are don't-care.  The
// thresholds ascend, so the flags of a byte are a thermometer code f_1 >= f_2
// >= ... >= f_7 and its rank t (0..7) is their count.
__device__ __forceinline__ void rank_flags(uint32_t w, const uint32_t (&nthr4)[7], uint32_t (&m)[7])
{
#pragma unroll
    for (int k = 0; k < 7; ++k) m[k] = avg4(nthr4[k], w);  // avg(~t, v) >= 128 <=> v >= t
}

// Three selector bits at bit 7 of each byte -> the 3-bit selector in bits
// 5..7 of the byte (bit 2 from s2); bits 0..4 are left as don't-care.
__device__ __forceinline__ uint32_t sel3(uint32_t s2, uint32_t s1, uint32_t s0)
{
    return bsel(0x80808080u, s2, bsel(0x40404040u, s1 >> 1, s0 >> 2));
}

// BC3 alpha block as two little-endian words, from the block's alpha rows.
__device__ __forceinline__ uint2 bc3_alpha(const AlphaRows& a)
{
    // level k = mn + floor(k*range/7); (x * 9363) >> 16 == x / 7 for x <= 1785.
    // Inverted: 255 - level = nmn - floor(k*r9 / 2^16) = the high half of
    // (nmn << 16 | 0xffff) - k*r9 (the low half only absorbs the borrow), so
    // one multiply-add and one v_perm_b32 of byte 2 give the replicated level.
    const uint32_t range = a.mx - a.mn, r9 = range * 9363u, nmn = 255u - a.mn;
    const uint32_t c = (nmn << 16) | 0xFFFFu;
    uint32_t nlv4[8];
    nlv4[0] = rep4(nmn);
#pragma unroll
    for (int k = 1; k < 7; ++k) nlv4[k] = perm(0u, c - (uint32_t)k * r9, 0x02020202u);
    nlv4[7] = rep4(255u - a.mx);
    uint32_t nthr4[7];
    rep_midpoints(nlv4, nthr4);
    uint32_t g[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        uint32_t m[7];
        rank_flags(a.w[r], nthr4, m);
        // rank t -> selector: 0 -> 1 (a1), 7 -> 0 (a0), 8 - t in between:
        //   t    0 1 2 3 4 5 6 7
        //   s2   0 1 1 1 1 0 0 0   = f1 & ~f5
        //   s1   0 1 1 0 0 1 1 0   = f1 ^ f3 ^ f5 ^ f7
        //   s0   1 1 0 1 0 1 0 0   = (f1 ^ ... ^ f6) | ~f1
        const uint32_t x = xor3(m[0], m[2], m[4]);
        const uint32_t s2 = m[0] & ~m[4], s1 = x ^ m[6];
        const uint32_t s0 = (x ^ xor3(m[1], m[3], m[5])) | ~m[0];
        const uint32_t z = sel3(s2, s1, s0);
        // columns 0, 1 -> bits 5..10 and 2, 3 -> bits 21..26, then all four
        // -> bits 5..16, column c at 5 + 3c (other bits don't-care)
        const uint32_t y = bsel(0x00E000E0u, z, z >> 5);
        g[r] = bsel(0x000007E0u, y, y >> 10);
    }
    // row r at bits 12r..12r+11 of the 48 index bits after a0, a1: word 0
    // takes row 0 and the low 4 bits of row 1, word 1 the rest
    const uint32_t head = a.mx | (a.mn << 8);
    const uint32_t w0 = bsel(0xFFFFu, head, bsel(0x0FFF0000u, g[0] << 11, g[1] << 23));
    const uint32_t w1 = bsel(0xFFu, g[1] >> 9, bsel(0xFFF00u, g[2] << 3, g[3] << 15));
    return make_uint2(w0, w1);
}

// EAC (ETC2 RGBA8 alpha) block as two little-endian words, from the block's
// alpha rows.
__device__ __forceinline__ uint2 eac_alpha(const AlphaRows& a)
{
    const uint32_t range = a.mx - a.mn;
    uint32_t m = ((range + 18u) * 3450u) >> 16;  // ceil(range / 19) for range <= 255
    m = m > 1u ? m : 1u;
    int32_t b = (int32_t)(a.mn + 10u * m);
    if (b > 255) b = (int32_t)a.mx - 9 * (int32_t)m;
    // Inverted levels 255 - clamp(b + m*s) = clamp(nb - m*s), nb = 255 - b,
    // two to a word in 16-bit lanes.  0 <= b <= 255 (the second base only
    // occurs for m <= 2 and mx > 235), so for the negative modifiers
    // nb - m*s >= 0 and only 255 bounds it, and for the positive ones
    // nb - m*s <= 255 and only 0 does: a v_pk_mad_u16 and one v_pk_min_u16,
    // or a v_pk_mul_lo_u16 and a saturating v_pk_sub_u16, per pair.
    const unsigned short nb = (unsigned short)(255 - b), ms = (unsigned short)m;
    const u16x2 nb2 = {nb, nb}, m2 = {ms, ms}, top = {255, 255};
    const u16x2 n01 = __builtin_elementwise_min(m2 * u16x2{10, 8} + nb2, top);  // modifiers -10, -8
    const u16x2 n23 = __builtin_elementwise_min(m2 * u16x2{5, 2} + nb2, top);   // -5, -2
    const u16x2 n45 = __builtin_elementwise_sub_sat(nb2, m2 * u16x2{1, 4});      // 1, 4
    const u16x2 n67 = __builtin_elementwise_sub_sat(nb2, m2 * u16x2{7, 9});      // 7, 9
    const uint32_t w[4] = {__builtin_bit_cast(uint32_t, n01), __builtin_bit_cast(uint32_t, n23),
                           __builtin_bit_cast(uint32_t, n45), __builtin_bit_cast(uint32_t, n67)};
    uint32_t nlv4[8];  // level k's byte in all four bytes
#pragma unroll
    for (int k = 0; k < 8; ++k) nlv4[k] = perm(0u, w[k >> 1], (k & 1) ? 0x02020202u : 0u);
    uint32_t nthr4[7];
    rep_midpoints(nlv4, nthr4);
    // Selector column groups: even columns (0, 2) and odd ones (1, 3) in the
    // two 16-bit halves, 12 bits each, row 0 in the top three.
    uint32_t ge = 0, go = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        uint32_t m[7];
        rank_flags(a.w[r], nthr4, m);
        // rank t -> selector: 0..3 -> 3..0 (negative modifiers), 4..7
        // unchanged: s2 = f4, s1 = ~(f2 ^ f6), s0 = ~(f1 ^ f2 ^ f3 ^ f5 ^ f6 ^ f7)
        const uint32_t s1 = ~(m[1] ^ m[5]);
        const uint32_t s0 = ~(xor3(m[0], m[1], m[2]) ^ xor3(m[4], m[5], m[6]));
        const uint32_t z = sel3(m[3], s1, s0);
        // column 0 / 2 selector at bits 5..7 / 21..23 -> 9-3r / 25-3r;
        // column 1 / 3 at bits 13..15 / 29..31 -> 9-3r / 25-3r
        const uint32_t dst = 0x0E000E00u >> (3 * r);
        ge = bsel(dst, r < 2 ? z << (4 - 3 * r) : z >> (3 * r - 4), ge);
        go = bsel(dst, z >> (4 + 3 * r), go);
    }
    // 48 bits, big-endian: column 0 (x = 0) first, 12 bits per column
    const uint32_t hi = ((ge & 0xFFFu) << 12) | (go & 0xFFFu);  // columns 0, 1
    const uint32_t lo = ((ge >> 16) << 12) | (go >> 16);        // columns 2, 3
    const uint32_t head = (uint32_t)b | (((m << 4) | 9u) << 8);
    return make_uint2(perm(hi, head, 0x05060100u), perm(lo, hi, 0x04050600u));
}

// DXT5 / ETC2 RGBA8 block: {alpha words, colour words}.  One pass over the
// pixels gives the colour box and the alpha range (block_stats_alpha).
template <int FMT>
__device__ __forceinline__ uint4 encode_block16(const uint32_t (&p)[16])
{
    uint32_t amn, amx;
    const BlockStats s = block_stats_alpha(p, amn, amx);
    const AlphaRows a = alpha_rows(p, amn, amx);
    if constexpr (FMT == kDXT5) {
        const uint2 al = bc3_alpha(a), c = encode_dxt1(p, s);
        return make_uint4(al.x, al.y, c.x, c.y);
    } else {
        const uint2 al = eac_alpha(a), c = encode_etc1(p, s);
        return make_uint4(al.x, al.y, c.x, c.y);
    }
}

}  // namespace goofy_dev
