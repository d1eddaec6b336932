/*
 * goofy_oracle.c -- CPU restatement of the Goofy DXT1 / ETC1S block encoder.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle: only the
 * test suite, the smoke check and the CPU-baseline leg of the benchmark may
 * load it, and only as the checker.  The product library never links it and
 * has no CPU fallback.
 *
 * It restates, one scalar block at a time, the 16-lane SSE2 pipeline of
 * GoofyTC/goofy_tc.h (goofySimdEncode<GOOFY_DXT1|GOOFY_ETC1>, :1069-1494) and
 * its two drivers (compressDXT1 :1497-1526, compressETC1 :1528-1557).  The
 * byte-lane semantics follow the generic backend of the same header
 * (avg :645-654, addsat :719-730, subsat :733-744, cmplti :708-716) which is
 * equivalent to the SSE2 backend (:168-396).
 *
 * Pinning (what checks this file, so that a wrong term fails a test):
 *   - bit-exact against the reference header itself, compiled from its own
 *     source into oracle/_ref/ (fuzzed blocks, every range value, special
 *     blocks, whole seeded images), and against the golden vectors under
 *     tests/golden/ that were produced by that build;
 *   - closed forms: constant / black / white blocks, q5(c) = c ? (c-1)>>3 : 0,
 *     the intensity-table breakpoints, c0 > c1 for every DXT1 block;
 *   - decode round trips with an independent BC1 / ETC1 decoder
 *     (oracle/goofy_decode_oracle.c): every decoded colour lies in the
 *     endpoint / base-colour palette the encoder chose.
 */
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#include "goofy_oracle.h"

/* ---- byte-lane primitives (goofy_tc.h generic backend) ------------------ */

/* _mm_avg_epu8: rounding average, goofy_tc.h:645-654 */
static inline uint8_t o_avg(uint8_t a, uint8_t b) { return (uint8_t)(((unsigned)a + b + 1u) >> 1); }
/* _mm_subs_epu8, goofy_tc.h:733-744 */
static inline uint8_t o_subsat(uint8_t a, uint8_t b) { return a > b ? (uint8_t)(a - b) : 0; }
/* _mm_adds_epu8, goofy_tc.h:719-730 */
static inline uint8_t o_addsat(uint8_t a, uint8_t b) { unsigned s = (unsigned)a + b; return s > 255u ? 255 : (uint8_t)s; }
static inline uint8_t o_min(uint8_t a, uint8_t b) { return a < b ? a : b; }
static inline uint8_t o_max(uint8_t a, uint8_t b) { return a > b ? a : b; }

/* YCoCg luma with the header's rounding order Y = avg(avg(R,B),G),
 * goofy_tc.h:1166 (endpoints) and :1205 (pixels). */
static inline uint8_t o_luma(uint8_t r, uint8_t g, uint8_t b) { return o_avg(o_avg(r, b), g); }

/* 8-bit -> 5-bit quantiser written as the header writes it: subtract eight
 * with saturation, then three rounding halvings, goofy_tc.h:1309,1311,1462. */
static inline uint8_t o_q5(uint8_t c)
{
    return o_avg(o_avg(o_avg(o_subsat(c, 8), 0), 0), 0);
}

/* Brightness range -> ETC1 control byte (byte 3 of the block): codeword
 * table index cw in both sub-block fields, diff bit and flip bit set.
 * Breakpoints of etc1BrighnessRangeTocontrolByte, goofy_tc.h:1040-1056. */
uint8_t goofy_oracle_etc1_control(uint8_t range)
{
    unsigned cw;
    if (range <= 21) cw = 0;
    else if (range <= 43) cw = 1;
    else if (range <= 73) cw = 2;
    else if (range <= 105) cw = 3;
    else if (range <= 151) cw = 4;
    else if (range <= 181) cw = 5;
    else if (range <= 253) cw = 6;
    else cw = 7;
    return (uint8_t)((cw << 5) | (cw << 2) | 0x3u);
}

static inline void put_le32(uint8_t* p, uint32_t v)
{
    p[0] = (uint8_t)v; p[1] = (uint8_t)(v >> 8); p[2] = (uint8_t)(v >> 16); p[3] = (uint8_t)(v >> 24);
}

/* ---- one 4x4 block ------------------------------------------------------- */

void goofy_oracle_analyse_block(const uint8_t* px, goofy_oracle_block_info* info)
{
    /* bounding box, goofy_tc.h:1103-1140 (alpha included, never used later) */
    uint8_t mn[4] = {255, 255, 255, 255}, mx[4] = {0, 0, 0, 0};
    for (int p = 0; p < 16; ++p)
        for (int c = 0; c < 4; ++c) {
            mn[c] = o_min(mn[c], px[4 * p + c]);
            mx[c] = o_max(mx[c], px[4 * p + c]);
        }
    /* endpoint brightness, range, mid, threshold, goofy_tc.h:1151-1190 */
    const uint8_t ymn = o_luma(mn[0], mn[1], mn[2]);
    const uint8_t ymx = o_luma(mx[0], mx[1], mx[2]);
    const uint8_t range = o_max(o_subsat(ymx, ymn), 8);
    const uint8_t mid = o_avg(ymn, ymx);
    const uint8_t half = o_avg(range, 0);
    const uint8_t quarter = o_avg(half, 0);
    const uint8_t eighth = o_avg(quarter, 0);
    const uint8_t qt = o_addsat(quarter, eighth);

    uint16_t gez = 0, lqt = 0;
    /* per-pixel projection and quantisation, goofy_tc.h:1196-1272 */
    for (int p = 0; p < 16; ++p) {
        const uint8_t y = o_luma(px[4 * p + 0], px[4 * p + 1], px[4 * p + 2]);
        const uint8_t pos = o_min(o_subsat(y, mid), 127);
        const uint8_t neg = o_min(o_subsat(mid, y), 127);
        const uint8_t absd = (uint8_t)(pos | neg);
        if (neg == 0) gez |= (uint16_t)(1u << p);
        /* signed byte compare (_mm_cmplt_epi8); both operands are <= 127 */
        if ((int8_t)absd < (int8_t)qt) lqt |= (uint16_t)(1u << p);
    }
    memcpy(info->mn, mn, 4);
    memcpy(info->mx, mx, 4);
    info->ymn = ymn; info->ymx = ymx; info->range = range; info->mid = mid; info->qt = qt;
    info->gez = gez; info->lqt = lqt;
}

static void encode_colour(int fmt, const uint8_t* px, uint8_t* out)
{
    goofy_oracle_block_info bi;
    goofy_oracle_analyse_block(px, &bi);

    if (fmt == GOOFY_ORACLE_DXT1) {
        /* 2-bit indices: bit 2p = (Y < mid), bit 2p+1 = near threshold,
         * goofy_tc.h:1292-1301 (zipB16(~Gez, Lqt) then movemask). */
        uint32_t idx = 0;
        for (int p = 0; p < 16; ++p) {
            idx |= (uint32_t)(((bi.gez >> p) & 1u) ^ 1u) << (2 * p);
            idx |= (uint32_t)((bi.lqt >> p) & 1u) << (2 * p + 1);
        }
        /* RGB565 endpoints; green LSB of colour 0 forced to 1 so that
         * c0 > c1 (4-colour mode), goofy_tc.h:1309-1356. */
        const uint32_t c0 = 0x20u | ((uint32_t)o_q5(bi.mx[0]) << 11) | ((uint32_t)o_q5(bi.mx[1]) << 6) | o_q5(bi.mx[2]);
        const uint32_t c1 = ((uint32_t)o_q5(bi.mn[0]) << 11) | ((uint32_t)o_q5(bi.mn[1]) << 6) | o_q5(bi.mn[2]);
        put_le32(out, c0 | (c1 << 16));
        put_le32(out + 4, idx);
        return;
    }

    /* ETC1S.  Average colour with the header's tree order: rows first
     * (avg(avg(r0,r1),avg(r2,r3))) per column, then columns, goofy_tc.h:1402-1414. */
    uint8_t colavg[4][4];
    for (int x = 0; x < 4; ++x)
        for (int c = 0; c < 4; ++c)
            colavg[x][c] = o_avg(o_avg(px[4 * (0 + x) + c], px[4 * (4 + x) + c]),
                                 o_avg(px[4 * (8 + x) + c], px[4 * (12 + x) + c]));
    uint8_t avgc[4];
    for (int c = 0; c < 4; ++c)
        avgc[c] = o_avg(o_avg(colavg[0][c], colavg[1][c]), o_avg(colavg[2][c], colavg[3][c]));

    /* Move the average colour's brightness onto mid, goofy_tc.h:1431-1449. */
    const uint8_t ya = o_luma(avgc[0], avgc[1], avgc[2]);
    const uint8_t pcorr = o_min(o_subsat(bi.mid, ya), 127);
    const uint8_t ncorr = o_min(o_subsat(ya, bi.mid), 127);
    const uint8_t corr = (uint8_t)(pcorr | ncorr);
    uint8_t base[3];
    for (int c = 0; c < 3; ++c)
        base[c] = (ncorr == 0) ? o_addsat(avgc[c], corr) : o_subsat(avgc[c], corr);

    /* Bytes 0..2: R5,G5,B5 in the high five bits with zero deltas
     * (differential mode); byte 3: table index, diff and flip bits,
     * goofy_tc.h:1462-1490. */
    const uint32_t w0 = ((uint32_t)goofy_oracle_etc1_control(bi.range) << 24) |
                        ((uint32_t)o_q5(base[0]) << 3) | ((uint32_t)o_q5(base[1]) << 11) |
                        ((uint32_t)o_q5(base[2]) << 19);

    /* Selectors, goofy_tc.h:1361-1395,1479: word bit j (0..15) belongs to the
     * pixel in row (j & 3) and column {2,3,0,1}[j >> 2] (transposeAs4x4x4);
     * the low half holds ~Gez and the high half ~Lqt.  Read as the standard
     * big-endian ETC1 selector planes this is MSB(k) = (Y < mid) and
     * LSB(k) = !near for pixel k = 4*x + y. */
    uint32_t gezbits = 0, lqtbits = 0;
    for (int j = 0; j < 16; ++j) {
        const int row = j & 3;
        const int col = ((j >> 2) + 2) & 3;
        const int p = 4 * row + col;
        gezbits |= (uint32_t)((bi.gez >> p) & 1u) << j;
        lqtbits |= (uint32_t)((bi.lqt >> p) & 1u) << j;
    }
    const uint32_t w1 = ~(gezbits | (lqtbits << 16));
    put_le32(out, w0);
    put_le32(out + 4, w1);
}

/* ---- alpha blocks (DXT5 / ETC2 RGBA8) ------------------------------------- */

static void alpha_range(const uint8_t* px, uint8_t a[16], int* amin, int* amax)
{
    *amin = 255;
    *amax = 0;
    for (int p = 0; p < 16; ++p) {
        a[p] = px[4 * p + 3];
        if (a[p] < *amin) *amin = a[p];
        if (a[p] > *amax) *amax = a[p];
    }
}

/* Rank of alpha value v among the sorted palette lv[0..7]: the number of
 * midpoints ceil((lv[k-1] + lv[k]) / 2), k = 1..7, that v reaches. */
static int nearest_rank(const int lv[8], int v)
{
    int t = 0;
    for (int k = 1; k < 8; ++k) t += v >= ((lv[k - 1] + lv[k] + 1) >> 1);
    return t;
}

void goofy_oracle_alpha_bc3(const uint8_t* px, uint8_t* out)
{
    uint8_t a[16];
    int amin, amax;
    alpha_range(px, a, &amin, &amax);
    const int range = amax - amin;
    /* Palette as the BC3 decoder forms it with a0 = amax > a1 = amin
     * (Src/decoder.cpp:874-896): entry 0 = a0, 1 = a1, 1+i = ((7-i) a0 +
     * i a1) / 7 truncated.  In ascending order, rank k holds amin +
     * floor(k * range / 7), which is entry 1 for k = 0, 8 - k for k = 1..6 and
     * 0 for k = 7. */
    int lv[8];
    for (int k = 0; k < 8; ++k) lv[k] = amin + (k * range) / 7;
    uint64_t bits = 0;
    for (int p = 0; p < 16; ++p) {
        const int t = nearest_rank(lv, a[p]);
        const int idx = t == 7 ? 0 : (t == 0 ? 1 : 8 - t);
        bits |= (uint64_t)idx << (3 * p); /* pixel p = 4*row + col, LSB first */
    }
    out[0] = (uint8_t)amax;
    out[1] = (uint8_t)amin;
    for (int i = 0; i < 6; ++i) out[2 + i] = (uint8_t)(bits >> (8 * i));
}

/* EAC modifier table 9 (the ETC2 alpha tables, Src/decoder.cpp:25-42), in
 * ascending order, and the selector of each rank. */
static const int kEac9Sorted[8] = {-10, -8, -5, -2, 1, 4, 7, 9};
static const int kEac9Index[8] = {3, 2, 1, 0, 4, 5, 6, 7};

void goofy_oracle_alpha_eac(const uint8_t* px, uint8_t* out)
{
    uint8_t a[16];
    int amin, amax;
    alpha_range(px, a, &amin, &amax);
    const int range = amax - amin;
    const int m = range ? (range + 18) / 19 : 1; /* 19 m >= range; m <= 14 */
    int b = amin + 10 * m;
    if (b > 255) b = amax - 9 * m; /* then b >= 0 (m = 14 needs amin <= 8) */
    int lv[8];
    for (int k = 0; k < 8; ++k) {
        const int v = b + m * kEac9Sorted[k];
        lv[k] = v < 0 ? 0 : (v > 255 ? 255 : v);
    }
    /* 48 selector bits, big-endian, pixel (x, y) at position 4x + y from the
     * top (Src/decoder.cpp:679-713). */
    uint64_t bits = 0;
    for (int x = 0; x < 4; ++x)
        for (int y = 0; y < 4; ++y) {
            const int idx = kEac9Index[nearest_rank(lv, a[4 * y + x])];
            bits |= (uint64_t)idx << (45 - 3 * (4 * x + y));
        }
    out[0] = (uint8_t)b;
    out[1] = (uint8_t)((m << 4) | 9);
    for (int i = 0; i < 6; ++i) out[2 + i] = (uint8_t)(bits >> (40 - 8 * i));
}

unsigned goofy_oracle_block_bytes(int fmt) { return fmt >= GOOFY_ORACLE_DXT5 ? 16u : 8u; }

void goofy_oracle_encode_block(int fmt, const uint8_t* px, uint8_t* out)
{
    if (fmt == GOOFY_ORACLE_DXT5) {
        goofy_oracle_alpha_bc3(px, out);
        encode_colour(GOOFY_ORACLE_DXT1, px, out + 8);
    } else if (fmt == GOOFY_ORACLE_ETC2) {
        goofy_oracle_alpha_eac(px, out);
        encode_colour(GOOFY_ORACLE_ETC1, px, out + 8);
    } else {
        encode_colour(fmt, px, out);
    }
}

/* ---- image drivers ------------------------------------------------------- */

static void gather_block(const uint8_t* src, size_t stride, unsigned bx, unsigned by, uint8_t* px)
{
    const uint8_t* base = src + (size_t)by * 4 * stride + (size_t)bx * 16;
    for (int r = 0; r < 4; ++r) memcpy(px + 16 * r, base + (size_t)r * stride, 16);
}

int goofy_oracle_compress_blocks(int fmt, uint8_t* dst, const uint8_t* src, unsigned width,
                                 unsigned height, size_t stride, unsigned row0, unsigned row1)
{
    if (width % 4 != 0) return -1;
    if (height % 4 != 0) return -2;
    const unsigned bw = width / 4;
    const size_t bb = goofy_oracle_block_bytes(fmt);
    uint8_t px[64];
    for (unsigned by = row0; by < row1; ++by)
        for (unsigned bx = 0; bx < bw; ++bx) {
            gather_block(src, stride, bx, by, px);
            goofy_oracle_encode_block(fmt, px, dst + ((size_t)by * bw + bx) * bb);
        }
    return 0;
}

/* Drop-in semantics of goofy::compressDXT1 / compressETC1, goofy_tc.h:1497-1557:
 * width % 16 -> -1 first, then height % 4 -> -2; output tightly packed. */
int goofy_oracle_compress(int fmt, uint8_t* dst, const uint8_t* src, unsigned width, unsigned height,
                          unsigned stride)
{
    if (width % 16 != 0) return -1;
    if (height % 4 != 0) return -2;
    return goofy_oracle_compress_blocks(fmt, dst, src, width, height, stride, 0, height / 4);
}
