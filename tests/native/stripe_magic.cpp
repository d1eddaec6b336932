// Host check of the launch-splitting arithmetic in internal.h: for block-row
// widths bw (powers of two and their neighbours up to 2^20, where the 40-bit
// division magic is inexact), every stripe of max_stripe_brows(bw) rows is a
// launch the kernels index correctly -- block count below 2^32, magic_ok, and
// (i * div_magic(bw)) >> 40 == i / bw at every row end and row start of the
// stripe.  Also the quality stripes (quality_stripe_rows) stay within one
// launch's 32-bit lane sums.  Prints "stripe magic ok" and exits 0.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "internal.h"

using namespace goofy_impl;

static int failures = 0;

static void check(bool ok, const char* what, uint64_t bw, uint64_t i)
{
    if (!ok && failures++ < 20) std::printf("FAIL %s: bw=%llu i=%llu\n", what, (unsigned long long)bw, (unsigned long long)i);
}

static void check_width(uint64_t bw)
{
    const uint32_t rows = max_stripe_brows(bw);
    if (bw >= (1ull << 20)) {
        check(rows == 0, "rows of 2^20 blocks or more are refused", bw, 0);
        return;
    }
    check(rows >= 1, "at least one row per stripe", bw, 0);
    const uint64_t nb = (uint64_t)rows * bw;
    check(nb <= 0xFFFFFFFFull, "block index fits 32 bits", bw, nb);
    check(magic_ok(bw, nb), "magic_ok for a full stripe", bw, nb);
    const uint64_t m = div_magic((uint32_t)bw);
    if (nb <= (1ull << 22)) {  // every block of the stripe
        for (uint64_t i = 0; i < nb; ++i) check(((i * m) >> 40) == i / bw, "row of block", bw, i);
    } else {  // the row ends and starts of sampled rows
        for (uint64_t r = 0; r < rows; r += (rows > 4096 ? rows / 4096 : 1))
            for (uint64_t i : {r * bw, r * bw + bw - 1}) check(((i * m) >> 40) == i / bw, "row of block", bw, i);
    }
    const uint64_t last = nb - 1;
    check(((last * m) >> 40) == last / bw, "row of the stripe's last block", bw, last);
    const uint32_t q = quality_stripe_rows(bw);
    check(q >= 1 && (uint64_t)q * bw <= kQualityMaxLaunchUnits && q <= rows, "quality stripe", bw, q);
}

int main()
{
    std::vector<uint64_t> widths;
    for (int k = 0; k <= 21; ++k) {
        const uint64_t p = 1ull << k;
        for (int64_t d = -3; d <= 3; ++d)
            if ((int64_t)p + d >= 1) widths.push_back(p + d);
    }
    for (uint64_t bw = (1ull << 19) - 64; bw <= (1ull << 19) + 64; ++bw) widths.push_back(bw);
    for (uint64_t bw = (1ull << 20) - 64; bw <= (1ull << 20); ++bw) widths.push_back(bw);
    for (uint64_t bw : {3ull, 7ull, 1000ull, 1023ull, 4097ull, 65535ull, 99999ull, 262143ull, 524287ull, 700001ull,
                        1048573ull, 1048575ull})
        widths.push_back(bw);
    for (uint64_t bw : widths) check_width(bw);
    // the smallest texture the per-texture batch path once launched unstriped
    // with an inexact magic: 1,048,575 blocks per row, 4 rows
    check(max_stripe_brows(1048575) == 1, "one row per stripe at 1,048,575 blocks per row", 1048575, 0);
    check(!magic_ok(1048575, 4ull * 1048575), "4 rows of 1,048,575 blocks exceed one launch", 1048575, 4);
    // and there the magic does give a wrong row (block 2,097,149 is in row 1)
    check(((2097149ull * div_magic(1048575)) >> 40) == 2, "the inexact magic is detectable", 1048575, 2097149);
    if (failures) {
        std::printf("%d failures\n", failures);
        return 1;
    }
    std::printf("stripe magic ok (%zu widths)\n", widths.size());
    return 0;
}
