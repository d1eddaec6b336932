/*
 * goofy_decode_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Independent BC1 and ETC1 block decoders (and BC3 / EAC alpha, decoded as
 * the harness's decoder does, Src/decoder.cpp:679-713, :874-925), written
 * from the public format definitions (BC1: two RGB565 endpoints, 2-bit indices LSB-first per pixel
 * in row-major order; ETC1: 64-bit big-endian word, individual/differential
 * base colours, eight intensity tables, MSB/LSB selector planes in
 * column-major pixel order; ETC2's T, H and planar colour modes, which a
 * differential sum outside 0..31 selects).  The rounding of the BC1 interpolants follows
 * the decoder the reference's quality harness uses
 * (Src/decoder.cpp:819-871: expand 5/6-bit by bit replication, thirds by
 * truncating division), and the ETC1 tables follow Src/decoder.cpp:388-410.
 *
 * Used for round-trip invariants on the encoder (every decoded colour lies in
 * the palette the encoder picked) and as the oracle of the GPU decoders.
 * Also holds the harness's RGB565 quality floor and the multi-threaded
 * striped driver of the encoder oracle.
 */
#include <pthread.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#include "goofy_oracle.h"

static inline int clamp255(int v) { return v < 0 ? 0 : (v > 255 ? 255 : v); }

/* four_only: BC3's colour block, always in 4-colour mode
 * (Src/decoder.cpp:945-956 decodes it with isDxt1 = false). */
static void decode_bc1_mode(const uint8_t* block, uint8_t* out, int four_only)
{
    const unsigned c0 = (unsigned)block[0] | ((unsigned)block[1] << 8);
    const unsigned c1 = (unsigned)block[2] | ((unsigned)block[3] << 8);
    int pal[4][4];
    const unsigned ep[2] = {c0, c1};
    for (int e = 0; e < 2; ++e) {
        const unsigned r = (ep[e] >> 11) & 31, g = (ep[e] >> 5) & 63, b = ep[e] & 31;
        pal[e][0] = (int)((r << 3) | (r >> 2));
        pal[e][1] = (int)((g << 2) | (g >> 4));
        pal[e][2] = (int)((b << 3) | (b >> 2));
        pal[e][3] = 255;
    }
    const int four = four_only || c0 > c1;
    for (int c = 0; c < 3; ++c) {
        if (four) {
            pal[2][c] = (2 * pal[0][c] + pal[1][c]) / 3;
            pal[3][c] = (pal[0][c] + 2 * pal[1][c]) / 3;
        } else {
            pal[2][c] = (pal[0][c] + pal[1][c]) / 2;
            pal[3][c] = 0;
        }
    }
    pal[2][3] = 255;
    pal[3][3] = four ? 255 : 0;
    const uint32_t idx = (uint32_t)block[4] | ((uint32_t)block[5] << 8) | ((uint32_t)block[6] << 16) |
                         ((uint32_t)block[7] << 24);
    for (int p = 0; p < 16; ++p) {
        const int k = (int)((idx >> (2 * p)) & 3u);
        for (int c = 0; c < 4; ++c) out[4 * p + c] = (uint8_t)pal[k][c];
    }
}

static const int kEtc1Tables[8][2] = {{2, 8}, {5, 17}, {9, 29}, {13, 42}, {18, 60}, {24, 80}, {33, 106}, {47, 183}};

/* ETC2 T and H modes: four paint colours, pixel (x, y) takes colour
 * 2*MSB + LSB of its selector bits k = 4x + y (no sign/magnitude mapping).
 * Both base colours are 4 bits per channel, widened by replication; the
 * distance table is 3, 6, 11, 16, 23, 32, 41, 64. */
static const int kEtc2Distance[8] = {3, 6, 11, 16, 23, 32, 41, 64};

static void etc2_paint(uint32_t lo, const int paint[4][3], uint8_t* out)
{
    for (int y = 0; y < 4; ++y)
        for (int x = 0; x < 4; ++x) {
            const int k = 4 * x + y;
            const int sel = (int)((((lo >> (16 + k)) & 1u) << 1) | ((lo >> k) & 1u));
            uint8_t* o = out + 4 * (4 * y + x);
            for (int c = 0; c < 3; ++c) o[c] = (uint8_t)paint[sel][c];
            o[3] = 255;
        }
}

static int widen4(int v) { return (v << 4) | v; }

/* T mode (the red differential sum leaves 0..31).  Fields of the big-endian
 * 64-bit block, bit 63 first: colour A red = bits 60-59 (high half) and 57-56,
 * green 55-52, blue 51-48; colour B red 47-44, green 43-40, blue 39-36;
 * distance index = bits 35-34 then bit 32.  Paint colours: A, B + d, B, B - d. */
static void etc2_t_mode(uint32_t hi, uint32_t lo, uint8_t* out)
{
    const int a[3] = {widen4((int)((((hi >> 27) & 3u) << 2) | ((hi >> 24) & 3u))), widen4((int)((hi >> 20) & 15u)),
                      widen4((int)((hi >> 16) & 15u))};
    const int b[3] = {widen4((int)((hi >> 12) & 15u)), widen4((int)((hi >> 8) & 15u)), widen4((int)((hi >> 4) & 15u))};
    const int d = kEtc2Distance[(((hi >> 2) & 3u) << 1) | (hi & 1u)];
    int paint[4][3];
    for (int c = 0; c < 3; ++c) {
        paint[0][c] = a[c];
        paint[1][c] = clamp255(b[c] + d);
        paint[2][c] = b[c];
        paint[3][c] = clamp255(b[c] - d);
    }
    etc2_paint(lo, paint, out);
}

/* H mode (red in range, the green sum leaves 0..31).  Colour A red = bits
 * 62-59, green = 58-56 then 52, blue = 51 then 49-47; colour B red 46-43,
 * green 42-39, blue 38-35.  Distance index = bit 34, bit 32, then 1 if A >= B
 * as 12-bit RGB444 numbers.  Paint colours: A + d, A - d, B + d, B - d. */
static void etc2_h_mode(uint32_t hi, uint32_t lo, uint8_t* out)
{
    const int a4[3] = {(int)((hi >> 27) & 15u), (int)((((hi >> 24) & 7u) << 1) | ((hi >> 20) & 1u)),
                       (int)((((hi >> 19) & 1u) << 3) | ((hi >> 15) & 7u))};
    const int b4[3] = {(int)((hi >> 11) & 15u), (int)((hi >> 7) & 15u), (int)((hi >> 3) & 15u)};
    const int va = (a4[0] << 8) | (a4[1] << 4) | a4[2], vb = (b4[0] << 8) | (b4[1] << 4) | b4[2];
    const int d = kEtc2Distance[(((hi >> 2) & 1u) << 2) | ((hi & 1u) << 1) | (va >= vb ? 1 : 0)];
    int paint[4][3];
    for (int c = 0; c < 3; ++c) {
        paint[0][c] = clamp255(widen4(a4[c]) + d);
        paint[1][c] = clamp255(widen4(a4[c]) - d);
        paint[2][c] = clamp255(widen4(b4[c]) + d);
        paint[3][c] = clamp255(widen4(b4[c]) - d);
    }
    etc2_paint(lo, paint, out);
}

/* Planar mode (red and green in range, the blue sum leaves 0..31): three
 * RGB676 colours O, H, V; pixel (x, y) = clamp((x (H - O) + y (V - O) + 4 O + 2) >> 2).
 * O: red 62-57, green 56 then 54-49, blue 48 then 44-43 then 41-39; H: red
 * 38-34 then 32, green 31-25, blue 24-19; V: red 18-13, green 12-6, blue 5-0. */
static void etc2_planar(uint32_t hi, uint32_t lo, uint8_t* out)
{
    const unsigned o6[3] = {(hi >> 25) & 63u, (((hi >> 24) & 1u) << 6) | ((hi >> 17) & 63u),
                            (((hi >> 16) & 1u) << 5) | (((hi >> 11) & 3u) << 3) | ((hi >> 7) & 7u)};
    const unsigned h6[3] = {(((hi >> 2) & 31u) << 1) | (hi & 1u), lo >> 25, (lo >> 19) & 63u};
    const unsigned v6[3] = {(lo >> 13) & 63u, (lo >> 6) & 127u, lo & 63u};
    int o[3], h[3], v[3];
    for (int c = 0; c < 3; ++c) {
        /* green has 7 bits, red and blue 6; widen by replicating the top bits */
        const int bits = c == 1 ? 7 : 6;
        o[c] = (int)((o6[c] << (8 - bits)) | (o6[c] >> (2 * bits - 8)));
        h[c] = (int)((h6[c] << (8 - bits)) | (h6[c] >> (2 * bits - 8)));
        v[c] = (int)((v6[c] << (8 - bits)) | (v6[c] >> (2 * bits - 8)));
    }
    for (int y = 0; y < 4; ++y)
        for (int x = 0; x < 4; ++x) {
            uint8_t* px = out + 4 * (4 * y + x);
            for (int c = 0; c < 3; ++c) {
                const int t = x * (h[c] - o[c]) + y * (v[c] - o[c]) + 4 * o[c] + 2;
                /* floor division by 4 (t may be negative) */
                px[c] = (uint8_t)clamp255(t >= 0 ? t / 4 : -((-t + 3) / 4));
            }
            px[3] = 255;
        }
}

/* ETC1 individual and differential modes, and the three modes ETC2 signals
 * with a differential sum outside 0..31 (red: T, else green: H, else blue:
 * planar), as the harness's decoder dispatches them (Src/decoder.cpp:716-773). */
void goofy_oracle_decode_etc1(const uint8_t* block, uint8_t* out)
{
    const uint32_t hi = ((uint32_t)block[0] << 24) | ((uint32_t)block[1] << 16) | ((uint32_t)block[2] << 8) | block[3];
    const uint32_t lo = ((uint32_t)block[4] << 24) | ((uint32_t)block[5] << 16) | ((uint32_t)block[6] << 8) | block[7];
    const int diff = (int)((hi >> 1) & 1u);
    const int flip = (int)(hi & 1u);
    if (diff) {
        for (int c = 0; c < 3; ++c) {
            const unsigned field = (hi >> (24 - 8 * c)) & 0xFFu;
            const int d = (int)(field & 7u) - ((field & 4u) ? 8 : 0);
            const int sum = (int)(field >> 3) + d;
            if (sum < 0 || sum > 31) {
                if (c == 0) etc2_t_mode(hi, lo, out);
                else if (c == 1) etc2_h_mode(hi, lo, out);
                else etc2_planar(hi, lo, out);
                return;
            }
        }
    }
    int base[2][3];
    for (int c = 0; c < 3; ++c) {
        const int shift = 24 - 8 * c; /* R at bits 31..24, G 23..16, B 15..8 */
        const unsigned field = (hi >> shift) & 0xFFu;
        if (diff) {
            const int b5 = (int)(field >> 3);
            int d = (int)(field & 7u);
            if (d >= 4) d -= 8;
            const int b5b = b5 + d; /* 0..31 here: other sums took the ETC2 modes above */
            base[0][c] = (b5 << 3) | (b5 >> 2);
            base[1][c] = (b5b << 3) | (b5b >> 2);
        } else {
            const int a4 = (int)(field >> 4), b4 = (int)(field & 15u);
            base[0][c] = (a4 << 4) | a4;
            base[1][c] = (b4 << 4) | b4;
        }
    }
    const int cw[2] = {(int)((hi >> 5) & 7u), (int)((hi >> 2) & 7u)};
    for (int y = 0; y < 4; ++y)
        for (int x = 0; x < 4; ++x) {
            const int k = 4 * x + y;
            const int sub = flip ? (y >= 2) : (x >= 2);
            const int msb = (int)((lo >> (16 + k)) & 1u);
            const int lsb = (int)((lo >> k) & 1u);
            const int mag = kEtc1Tables[cw[sub]][lsb];
            const int delta = msb ? -mag : mag;
            uint8_t* o = out + 4 * (4 * y + x);
            for (int c = 0; c < 3; ++c) o[c] = (uint8_t)clamp255(base[sub][c] + delta);
            o[3] = 255;
        }
}

void goofy_oracle_decode_bc1(const uint8_t* block, uint8_t* out) { decode_bc1_mode(block, out, 0); }

/* BC3 alpha: a0 > a1 gives 8 entries, a0, a1 and six at sevenths; otherwise
 * 6 entries at fifths plus 0 and 255.  Interpolants truncate, as in the
 * harness's decoder (Src/decoder.cpp:874-896).  48 index bits follow,
 * little-endian, 3 per pixel in row-major order. */
static void decode_bc3_alpha(const uint8_t* block, uint8_t* out)
{
    const int a0 = block[0], a1 = block[1];
    int pal[8] = {a0, a1, 0, 0, 0, 0, 0, 0};
    if (a0 > a1) {
        for (int i = 1; i < 7; ++i) pal[1 + i] = ((7 - i) * a0 + i * a1) / 7;
    } else {
        for (int i = 1; i < 5; ++i) pal[1 + i] = ((5 - i) * a0 + i * a1) / 5;
        pal[6] = 0;
        pal[7] = 255;
    }
    uint64_t bits = 0;
    for (int i = 0; i < 6; ++i) bits |= (uint64_t)block[2 + i] << (8 * i);
    for (int p = 0; p < 16; ++p) out[4 * p + 3] = (uint8_t)pal[(bits >> (3 * p)) & 7u];
}

/* ETC2 alpha (EAC): base, multiplier (high nibble) and table (low nibble),
 * then 48 big-endian selector bits, pixel (x, y) at position 4x + y from the
 * top; alpha = clamp(base + table[sel] * multiplier). */
static const int kEacTables[16][8] = {
    {-3, -6, -9, -15, 2, 5, 8, 14}, {-3, -7, -10, -13, 2, 6, 9, 12}, {-2, -5, -8, -13, 1, 4, 7, 12},
    {-2, -4, -6, -13, 1, 3, 5, 12}, {-3, -6, -8, -12, 2, 5, 7, 11}, {-3, -7, -9, -11, 2, 6, 8, 10},
    {-4, -7, -8, -11, 3, 6, 7, 10}, {-3, -5, -8, -11, 2, 4, 7, 10}, {-2, -6, -8, -10, 1, 5, 7, 9},
    {-2, -5, -8, -10, 1, 4, 7, 9},  {-2, -4, -8, -10, 1, 3, 7, 9},  {-2, -5, -7, -10, 1, 4, 6, 9},
    {-3, -4, -7, -10, 2, 3, 6, 9},  {-1, -2, -3, -10, 0, 1, 2, 9},  {-4, -6, -8, -9, 3, 5, 7, 8},
    {-3, -5, -7, -9, 2, 4, 6, 8}};

static void decode_eac_alpha(const uint8_t* block, uint8_t* out)
{
    const int base = block[0], mult = block[1] >> 4, table = block[1] & 15;
    uint64_t bits = 0;
    for (int i = 0; i < 6; ++i) bits = (bits << 8) | block[2 + i];
    for (int x = 0; x < 4; ++x)
        for (int y = 0; y < 4; ++y) {
            const int sel = (int)((bits >> (45 - 3 * (4 * x + y))) & 7u);
            out[4 * (4 * y + x) + 3] = (uint8_t)clamp255(base + kEacTables[table][sel] * mult);
        }
}

void goofy_oracle_decode_bc3(const uint8_t* block, uint8_t* out)
{
    decode_bc1_mode(block + 8, out, 1);
    decode_bc3_alpha(block, out);
}

void goofy_oracle_decode_etc2(const uint8_t* block, uint8_t* out)
{
    goofy_oracle_decode_etc1(block + 8, out);
    decode_eac_alpha(block, out);
}

int goofy_oracle_decode_image(int fmt, uint8_t* dst, const uint8_t* blocks, unsigned width, unsigned height)
{
    if (width % 4 != 0) return -1;
    if (height % 4 != 0) return -2;
    const unsigned bw = width / 4, bh = height / 4;
    uint8_t px[64];
    for (unsigned by = 0; by < bh; ++by)
        for (unsigned bx = 0; bx < bw; ++bx) {
            const uint8_t* b = blocks + ((size_t)by * bw + bx) * goofy_oracle_block_bytes(fmt);
            if (fmt == GOOFY_ORACLE_DXT1) goofy_oracle_decode_bc1(b, px);
            else if (fmt == GOOFY_ORACLE_ETC1) goofy_oracle_decode_etc1(b, px);
            else if (fmt == GOOFY_ORACLE_DXT5) goofy_oracle_decode_bc3(b, px);
            else goofy_oracle_decode_etc2(b, px);
            for (int r = 0; r < 4; ++r)
                memcpy(dst + ((size_t)by * 4 + r) * width * 4 + (size_t)bx * 16, px + 16 * r, 16);
        }
    return 0;
}

/* ---- striped multi-threaded encoder driver ------------------------------ */

typedef struct {
    int fmt;
    uint8_t* dst;
    const uint8_t* src;
    unsigned width, height, row0, row1;
    size_t stride;
} stripe_job;

/* Src/main.cpp:501-555, written per channel from its arithmetic: a pixel is
 * a little-endian word whose byte 2 the harness calls "r", byte 1 "g" and
 * byte 0 "b".  Box average (sum of 4 + 2) / 4; "r" keeps 5 bits and is
 * widened r5*8 + r5/4; "g" keeps 6 bits, g6*4 + g6/16; "b" keeps 6 bits but
 * is widened like a 5-bit value, b6*4 OR (b6/4 mod 8), as the harness has it. */
void goofy_oracle_rgb565_baseline(uint8_t* dst, size_t dst_pitch, const uint8_t* src, size_t src_pitch,
                                  unsigned width, unsigned height)
{
    for (unsigned y = 0; y + 1 < height; y += 2) {
        for (unsigned x = 0; x + 1 < width; x += 2) {
            unsigned avg[3];
            for (int c = 0; c < 3; ++c) {
                const unsigned sum = src[y * src_pitch + 4 * x + c] + src[y * src_pitch + 4 * (x + 1) + c] +
                                     src[(y + 1) * src_pitch + 4 * x + c] +
                                     src[(y + 1) * src_pitch + 4 * (x + 1) + c];
                avg[c] = (sum + 2) / 4;
            }
            const unsigned b6 = avg[0] / 4, g6 = avg[1] / 4, r5 = avg[2] / 8;
            const uint8_t out[4] = {(uint8_t)((b6 * 4) | ((b6 / 4) % 8)), (uint8_t)(g6 * 4 + g6 / 16),
                                    (uint8_t)(r5 * 8 + r5 / 4), 255};
            for (unsigned dy = 0; dy < 2; ++dy)
                for (unsigned dx = 0; dx < 2; ++dx) memcpy(dst + (y + dy) * dst_pitch + 4 * (x + dx), out, 4);
        }
    }
}

static void* stripe_main(void* arg)
{
    const stripe_job* j = (const stripe_job*)arg;
    goofy_oracle_compress_blocks(j->fmt, j->dst, j->src, j->width, j->height, j->stride, j->row0, j->row1);
    return NULL;
}

int goofy_oracle_compress_mt(int fmt, uint8_t* dst, const uint8_t* src, unsigned width, unsigned height,
                             size_t stride, int nthreads)
{
    if (width % 4 != 0) return -1;
    if (height % 4 != 0) return -2;
    const unsigned bh = height / 4;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    if ((unsigned)nthreads > bh) nthreads = bh ? (int)bh : 1;
    pthread_t th[256];
    stripe_job jobs[256];
    for (int t = 0; t < nthreads; ++t) {
        jobs[t].fmt = fmt; jobs[t].dst = dst; jobs[t].src = src;
        jobs[t].width = width; jobs[t].height = height; jobs[t].stride = stride;
        jobs[t].row0 = (unsigned)((uint64_t)bh * t / nthreads);
        jobs[t].row1 = (unsigned)((uint64_t)bh * (t + 1) / nthreads);
    }
    int started = 0;
    for (int t = 1; t < nthreads; ++t) {
        if (pthread_create(&th[t], NULL, stripe_main, &jobs[t]) != 0) break;
        started = t;
    }
    stripe_main(&jobs[0]);
    for (int t = 1; t <= started; ++t) pthread_join(th[t], NULL);
    /* any stripe whose thread could not start runs here */
    for (int t = started + 1; t < nthreads; ++t) stripe_main(&jobs[t]);
    return 0;
}
