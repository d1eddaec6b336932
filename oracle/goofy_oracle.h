/*
 * goofy_oracle.h -- TEST INFRASTRUCTURE ONLY: the CPU oracle of the Goofy
 * DXT1 / ETC1S block encoder (restatement of GoofyTC/goofy_tc.h), the alpha
 * blocks of the DXT5 / ETC2 RGBA8 forms, and independent BC1 / ETC1 / BC3 /
 * EAC block decoders used for round-trip invariants.
 * Only tests, the smoke check and the benchmark's CPU-baseline leg use it.
 */
#ifndef GOOFY_ORACLE_H
#define GOOFY_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* DXT5 (BC3) and ETC2 RGBA8 (EAC) are 16-byte blocks: an 8-byte alpha block
 * followed by the DXT1 / ETC1S colour block of the same pixels. */
enum { GOOFY_ORACLE_DXT1 = 0, GOOFY_ORACLE_ETC1 = 1, GOOFY_ORACLE_DXT5 = 2, GOOFY_ORACLE_ETC2 = 3 };

typedef struct goofy_oracle_block_info {
    uint8_t mn[4], mx[4];      /* per-channel bounding box (RGBA) */
    uint8_t ymn, ymx;          /* luma of the two box corners */
    uint8_t range, mid, qt;    /* max(ymx-ymn,8), avg(ymn,ymx), ~0.375*range */
    uint16_t gez;              /* bit p: Y(p) >= mid   (p = 4*row + col) */
    uint16_t lqt;              /* bit p: |Y(p)-mid| < qt */
} goofy_oracle_block_info;

/* One block: px = 16 RGBA8 pixels, row-major (64 bytes); out = 8 bytes
 * (DXT1, ETC1S) or 16 bytes (DXT5, ETC2 RGBA8). */
void goofy_oracle_analyse_block(const uint8_t* px, goofy_oracle_block_info* info);
void goofy_oracle_encode_block(int fmt, const uint8_t* px, uint8_t* out);
unsigned goofy_oracle_block_bytes(int fmt);

/* The alpha blocks on their own (8 bytes).  The reference has no alpha
 * encoder (its README lists a DXT5 / ETC2 alpha encoder that works on alpha
 * directly instead of brightness as the next step); these restate the rule
 * the library's kernels implement:
 *   BC3: a0 = max alpha, a1 = min alpha (8-level mode whenever they differ);
 *   EAC: table 9, multiplier m = max(1, ceil((max-min)/19)), base
 *        b = min + 10m (lowest level on the minimum), or b = max - 9m (highest
 *        level on the maximum) when min + 10m > 255;
 *   both: every pixel takes the palette entry nearest to it, as the decoder
 *   reconstructs the palette; a pixel exactly between two entries takes the
 *   higher one. */
void goofy_oracle_alpha_bc3(const uint8_t* px, uint8_t* out);
void goofy_oracle_alpha_eac(const uint8_t* px, uint8_t* out);
uint8_t goofy_oracle_etc1_control(uint8_t range);

/* Drop-in semantics (width % 16, height % 4; returns 0 / -1 / -2). */
int goofy_oracle_compress(int fmt, uint8_t* dst, const uint8_t* src, unsigned width, unsigned height,
                          unsigned stride);
/* Per-block semantics for any width % 4 == 0, block rows [row0, row1). */
int goofy_oracle_compress_blocks(int fmt, uint8_t* dst, const uint8_t* src, unsigned width,
                                 unsigned height, size_t stride, unsigned row0, unsigned row1);
/* Multi-threaded per-block driver: one stripe of block rows per thread. */
int goofy_oracle_compress_mt(int fmt, uint8_t* dst, const uint8_t* src, unsigned width,
                             unsigned height, size_t stride, int nthreads);

/* Independent decoders (BC1 4-colour/3-colour; ETC1 individual and
 * differential modes and ETC2's T, H and planar colour modes).  out = 16
 * RGBA8 pixels, row-major. */
void goofy_oracle_decode_bc1(const uint8_t* block, uint8_t* out);
void goofy_oracle_decode_etc1(const uint8_t* block, uint8_t* out);
/* 16-byte blocks: BC3 (alpha block, then a colour block always decoded in
 * 4-colour mode) and ETC2 RGBA8 (EAC alpha block, then an ETC2 colour
 * block in any of its five modes). */
void goofy_oracle_decode_bc3(const uint8_t* block, uint8_t* out);
void goofy_oracle_decode_etc2(const uint8_t* block, uint8_t* out);
/* Whole images, tightly packed block rows; dst pitch = width*4. */
int goofy_oracle_decode_image(int fmt, uint8_t* dst, const uint8_t* blocks, unsigned width,
                              unsigned height);

/* The benchmark harness's quality floor (generateDownsamlpedRgb565Test,
 * Src/main.cpp:501-555): each 2x2 cell -> rounded box average, cut to 5/6/6
 * bits and widened back as the harness does; width and height even. */
void goofy_oracle_rgb565_baseline(uint8_t* dst, size_t dst_pitch, const uint8_t* src, size_t src_pitch,
                                  unsigned width, unsigned height);

#ifdef __cplusplus
}
#endif
#endif
