// The per-block arithmetic of the gfx950 kernels (goofy_amd/csrc/block_codec.h
// and block_decode.h) built for the host against tests/native/host_shim, so
// that CPU tests can compare every encoder and decoder with the oracle.
//   codec_host encode IN OUT      IN: n blocks of 64 bytes (16 RGBA8 pixels,
//                                 row-major); OUT: the DXT1, ETC1S, DXT5 and
//                                 ETC2 RGBA8 encodings, one format after the
//                                 other (8n + 8n + 16n + 16n bytes)
//   codec_host decode FMT IN OUT  IN: n blocks of format FMT (8 or 16 bytes);
//                                 OUT: n x 64 bytes of decoded pixels
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "block_codec.h"
#include "block_decode.h"

using namespace goofy_dev;

static std::vector<unsigned char> read_all(const char* path)
{
    std::vector<unsigned char> v;
    FILE* f = std::fopen(path, "rb");
    if (!f) std::exit(3);
    unsigned char buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) v.insert(v.end(), buf, buf + got);
    std::fclose(f);
    return v;
}

static void write_all(const char* path, const std::vector<unsigned char>& v)
{
    FILE* f = std::fopen(path, "wb");
    if (!f || std::fwrite(v.data(), 1, v.size(), f) != v.size()) std::exit(4);
    std::fclose(f);
}

template <int FMT>
static void decode_all(const std::vector<unsigned char>& in, std::vector<unsigned char>& out)
{
    const size_t bb = FMT >= 2 ? 16 : 8, n = in.size() / bb;
    out.resize(n * 64);
    for (size_t i = 0; i < n; ++i) {
        uint32_t w[4] = {0, 0, 0, 0};
        std::memcpy(w, in.data() + bb * i, bb);
        uint32_t px[16];
        decode_block<FMT>(make_uint4(w[0], w[1], w[2], w[3]), px);
        std::memcpy(out.data() + 64 * i, px, 64);
    }
}

int main(int argc, char** argv)
{
    if (argc == 4 && std::string(argv[1]) == "encode") {
        const std::vector<unsigned char> in = read_all(argv[2]);
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
        write_all(argv[3], out);
        return 0;
    }
    if (argc == 5 && std::string(argv[1]) == "decode") {
        const std::vector<unsigned char> in = read_all(argv[3]);
        std::vector<unsigned char> out;
        switch (std::atoi(argv[2])) {
        case 0: decode_all<0>(in, out); break;
        case 1: decode_all<1>(in, out); break;
        case 2: decode_all<2>(in, out); break;
        case 3: decode_all<3>(in, out); break;
        default: return 2;
        }
        write_all(argv[4], out);
        return 0;
    }
    return 2;
}
