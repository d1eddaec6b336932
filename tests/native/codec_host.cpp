// The per-block arithmetic of the gfx950 kernels (goofy_amd/csrc/block_codec.h)
// built for the host against tests/native/host_shim, so that a CPU test can
// compare every encoder with the oracle.  Usage: codec_host IN OUT, where IN
// holds n blocks of 64 bytes (16 RGBA8 pixels, row-major) and OUT receives the
// DXT1, ETC1S, DXT5 and ETC2 RGBA8 encodings, one format after the other
// (8n + 8n + 16n + 16n bytes).
#include <cstdio>
#include <cstring>
#include <vector>

#include "block_codec.h"

using namespace goofy_dev;

int main(int argc, char** argv)
{
    if (argc != 3) return 2;
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 3;
    std::vector<unsigned char> in;
    unsigned char buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) in.insert(in.end(), buf, buf + got);
    std::fclose(f);
    const size_t n = in.size() / 64;
    std::vector<unsigned char> out(n * 48);
    unsigned char* o8[2] = {out.data(), out.data() + 8 * n};
    unsigned char* o16[2] = {out.data() + 16 * n, out.data() + 32 * n};
    for (size_t i = 0; i < n; ++i) {
        uint32_t p[16];
        std::memcpy(p, in.data() + 64 * i, 64);
        const uint2 a = encode_block<kDXT1>(p), b = encode_block<kETC1>(p);
        const uint4 c = encode_block16<kDXT5>(p), d = encode_block16<kETC2>(p);
        std::memcpy(o8[0] + 8 * i, &a, 8);
        std::memcpy(o8[1] + 8 * i, &b, 8);
        std::memcpy(o16[0] + 16 * i, &c, 16);
        std::memcpy(o16[1] + 16 * i, &d, 16);
    }
    f = std::fopen(argv[2], "wb");
    if (!f || std::fwrite(out.data(), 1, out.size(), f) != out.size()) return 4;
    std::fclose(f);
    return 0;
}
