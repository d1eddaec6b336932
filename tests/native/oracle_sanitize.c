/* The oracle under AddressSanitizer + UndefinedBehaviorSanitizer (built by
 * tests/test_native_host.py, host code only): random blocks, odd widths with
 * padded strides through the block and multi-threaded drivers, and every
 * decoder mode including ETC1 differential sums outside 0..31. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "goofy_oracle.h"

static uint32_t rng_state = 12345u;
static uint32_t rnd(void)
{
    rng_state ^= rng_state << 13;
    rng_state ^= rng_state >> 17;
    rng_state ^= rng_state << 5;
    return rng_state;
}

int main(void)
{
    uint8_t px[64], out[8], dec[64];
    for (int i = 0; i < 200000; ++i) {
        for (int k = 0; k < 64; ++k) px[k] = (uint8_t)rnd();
        goofy_oracle_encode_block(i & 1, px, out);
        for (int k = 0; k < 8; ++k) out[k] = (uint8_t)rnd();
        goofy_oracle_decode_bc1(out, dec);
        goofy_oracle_decode_etc1(out, dec);
    }
    for (unsigned w = 4; w <= 68; w += 4)
        for (unsigned h = 4; h <= 12; h += 4) {
            const size_t stride = (size_t)w * 4 + (rnd() % 3) * 4;
            uint8_t* img = malloc(stride * h);
            uint8_t* a = malloc((w / 4) * (h / 4) * 8);
            uint8_t* b = malloc((w / 4) * (h / 4) * 8);
            uint8_t* rgba = malloc((size_t)w * h * 4);
            for (size_t k = 0; k < stride * h; ++k) img[k] = (uint8_t)rnd();
            if (goofy_oracle_compress_blocks(0, a, img, w, h, stride, 0, h / 4) != 0) return 1;
            if (goofy_oracle_compress_mt(0, b, img, w, h, stride, 3) != 0) return 2;
            if (memcmp(a, b, (w / 4) * (h / 4) * 8) != 0) return 3;
            if (goofy_oracle_decode_image(1, rgba, a, w, h) != 0) return 4;
            if (w % 16 == 0 && goofy_oracle_compress(1, a, img, w, h, (unsigned)stride) != 0) return 5;
            free(img); free(a); free(b); free(rgba);
        }
    printf("oracle sanitize ok\n");
    return 0;
}
