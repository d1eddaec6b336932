// A caller of the reference API (README.md:87-104 usage), unchanged except
// that it includes this build's include/goofy_tc.h and links libgoofy_hip.so.
// Encodes a seeded 512x512 image in both formats and checks the reference's
// return codes; writes the blocks to argv[1] (DXT1) and argv[2] (ETC1S).  Also
// calls the two 16-byte formats, which have no reference counterpart, on the
// same (opaque) image: their colour halves must be those blocks, and their
// alpha blocks the closed forms of an all-255 alpha block.
// Built by goofy_amd/build.py as tests/native/libdropin_example.so (the
// entry point is dropin_main) so that it ships with the built libraries.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "goofy_hip.h"  // only for the seeded test image
#include "goofy_tc.h"

extern "C" int dropin_main(int argc, char** argv)
{
    if (argc < 3) return 2;
    const unsigned w = 512, h = 512, stride = w * 4;
    std::vector<unsigned char> rgba(stride * h);
    if (goofy_synth_fill_host(0, 0x601F0001, 0, 12, rgba.data(), w, h, stride, 0) != 0) return 3;
    std::vector<unsigned char> dxt1((w / 4) * (h / 4) * 8), etc1(dxt1.size());
    // the reference's argument checks, in its order
    if (goofy::compressDXT1(dxt1.data(), rgba.data(), 500, h, stride) != -1) return 4;
    if (goofy::compressETC1(etc1.data(), rgba.data(), w, 510, stride) != -2) return 5;
    if (goofy::compressDXT1(dxt1.data(), rgba.data(), w, h, stride) != 0) return 6;
    if (goofy::compressETC1(etc1.data(), rgba.data(), w, h, stride) != 0) return 7;
    std::vector<unsigned char> dxt5(dxt1.size() * 2), etc2(dxt1.size() * 2);
    if (goofy::compressDXT5(dxt5.data(), rgba.data(), w, h, stride) != 0) return 10;
    if (goofy::compressETC2(etc2.data(), rgba.data(), w, h, stride) != 0) return 11;
    // BC3: a0 = a1 = 255, every index 0; EAC: base 255 - 9 = 246, multiplier 1,
    // table 9, every selector 7 (+9), big-endian bits all ones
    const unsigned char bc3_alpha[8] = {0xFF, 0xFF, 0, 0, 0, 0, 0, 0};
    const unsigned char eac_alpha[8] = {0xF6, 0x19, 0xFF, 0xFF, 0xFF, 0xFF, 0xFF, 0xFF};
    for (size_t i = 0; i < dxt1.size() / 8; ++i) {
        if (std::memcmp(&dxt5[16 * i], bc3_alpha, 8) || std::memcmp(&dxt5[16 * i + 8], &dxt1[8 * i], 8)) return 12;
        if (std::memcmp(&etc2[16 * i], eac_alpha, 8) || std::memcmp(&etc2[16 * i + 8], &etc1[8 * i], 8)) return 13;
    }
    FILE* a = std::fopen(argv[1], "wb");
    FILE* b = std::fopen(argv[2], "wb");
    if (!a || !b) return 8;
    std::fwrite(dxt1.data(), 1, dxt1.size(), a);
    std::fwrite(etc1.data(), 1, etc1.size(), b);
    std::fclose(a);
    std::fclose(b);
    std::printf("dropin_example ok\n");
    return 0;
}
