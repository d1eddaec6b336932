// containers.cpp -- DDS and KTX1 files around encoded blocks.
//
// The headers are the ones the reference's harness writes next to every
// encode (saveDds / saveKtx, Src/main.cpp:154-220; header structs :78-136):
//   DDS : "DDS " magic, 124-byte header with flags 0x000A1007, width, height,
//         depth 1, pitchOrLinearSize = payload bytes, one mip level, a FOURCC
//         pixel format ("DXT1" or "DXT5"), caps fields left zero as the
//         harness does;
//   KTX1: the 12-byte identifier, endianness 0x04030201, glTypeSize 1,
//         glInternalFormat GL_COMPRESSED_RGB_S3TC_DXT1_EXT (0x83F0),
//         GL_ETC1_RGB8_OES (0x8D64), GL_COMPRESSED_RGBA_S3TC_DXT5_EXT (0x83F3)
//         or GL_COMPRESSED_RGBA8_ETC2_EAC (0x9278) (Src/main.cpp:64-73), base
//         format GL_RGB for ETC1 and GL_RGBA otherwise, one face, one mip
//         level, no key/value data, then the 32-bit image size and the blocks.
// All fields little-endian.
#include <cstdio>
#include <cstring>

#include "goofy_hip.h"
#include "internal.h"

using namespace goofy_impl;

namespace {

void put32(unsigned char* p, uint32_t v)
{
    p[0] = (unsigned char)v;
    p[1] = (unsigned char)(v >> 8);
    p[2] = (unsigned char)(v >> 16);
    p[3] = (unsigned char)(v >> 24);
}

constexpr size_t kDdsHeader = 128;
constexpr size_t kKtxHeader = 64 + 4;

size_t header_bytes(int kind) { return kind == GOOFY_CONTAINER_DDS ? kDdsHeader : kKtxHeader; }

void write_header(int kind, int fmt, unsigned w, unsigned h, size_t payload, unsigned char* p)
{
    std::memset(p, 0, header_bytes(kind));
    if (kind == GOOFY_CONTAINER_DDS) {
        put32(p + 0, 0x20534444u);  // "DDS "
        put32(p + 4, 124);          // header size without the magic
        put32(p + 8, 0x000A1007u);  // CAPS | HEIGHT | WIDTH | PIXELFORMAT | MIPMAPCOUNT | LINEARSIZE
        put32(p + 12, h);
        put32(p + 16, w);
        put32(p + 20, (uint32_t)payload);
        put32(p + 24, 1);           // depth
        put32(p + 28, 1);           // mip levels
        put32(p + 76, 32);          // pixel format size
        put32(p + 80, 0x4);         // FOURCC
        put32(p + 84, fmt == GOOFY_FORMAT_DXT5 ? 0x35545844u : 0x31545844u);  // "DXT5" / "DXT1"
        return;
    }
    static const unsigned char id[12] = {0xAB, 0x4B, 0x54, 0x58, 0x20, 0x31, 0x31, 0xBB, 0x0D, 0x0A, 0x1A, 0x0A};
    std::memcpy(p, id, 12);
    put32(p + 12, 0x04030201u);  // endianness
    put32(p + 16, 0);            // glType (compressed)
    put32(p + 20, 1);            // glTypeSize
    put32(p + 24, 0);            // glFormat (compressed)
    static const uint32_t gl_format[4] = {0x83F0u, 0x8D64u, 0x83F3u, 0x9278u};
    put32(p + 28, gl_format[fmt]);
    put32(p + 32, fmt == GOOFY_FORMAT_ETC1 ? 0x1907u : 0x1908u);
    put32(p + 36, w);
    put32(p + 40, h);
    put32(p + 44, 0);  // depth
    put32(p + 48, 0);  // array elements
    put32(p + 52, 1);  // faces
    put32(p + 56, 1);  // mip levels
    put32(p + 60, 0);  // key/value bytes
    put32(p + 64, (uint32_t)payload);
}

int check(int kind, int fmt, unsigned w, unsigned h, size_t payload)
{
    if (kind != GOOFY_CONTAINER_DDS && kind != GOOFY_CONTAINER_KTX) return fail(GOOFY_ERR_INVALID, "unknown container");
    if (!format_valid(fmt)) return fail(GOOFY_ERR_INVALID, "unknown format");
    if (kind == GOOFY_CONTAINER_DDS && fmt != GOOFY_FORMAT_DXT1 && fmt != GOOFY_FORMAT_DXT5)
        return fail(GOOFY_ERR_INVALID, "DDS holds DXT1 and DXT5 only (ETC formats go in KTX)");
    if (w % 4 != 0) return fail(GOOFY_ERR_WIDTH, "width must be a multiple of 4");
    if (h % 4 != 0) return fail(GOOFY_ERR_HEIGHT, "height must be a multiple of 4");
    if (payload != (size_t)(w / 4) * (h / 4) * block_bytes(fmt))
        return fail(GOOFY_ERR_INVALID, "payload is not (w/4)*(h/4) blocks of the format's size");
    if (payload > 0xFFFFFFFFull) return fail(GOOFY_ERR_TOO_LARGE, "payload above 4 GiB");
    return GOOFY_OK;
}

}  // namespace

extern "C" {

size_t goofy_container_bytes(int kind, int format, unsigned int width, unsigned int height)
{
    return header_bytes(kind) + (size_t)(width / 4) * (height / 4) * (format_valid(format) ? block_bytes(format) : 8);
}

int goofy_container_write(int kind, int format, unsigned int width, unsigned int height, const void* blocks,
                          size_t blocks_bytes, void* out, size_t out_capacity, size_t* written)
{
    clear_error();
    int rc = check(kind, format, width, height, blocks_bytes);
    if (rc != GOOFY_OK) return rc;
    if (!out || (blocks_bytes && !blocks)) return fail(GOOFY_ERR_INVALID, "null buffer");
    const size_t total = header_bytes(kind) + blocks_bytes;
    if (out_capacity < total) return fail(GOOFY_ERR_INVALID, "output buffer too small");
    unsigned char* o = static_cast<unsigned char*>(out);
    write_header(kind, format, width, height, blocks_bytes, o);
    if (blocks_bytes) std::memcpy(o + header_bytes(kind), blocks, blocks_bytes);
    if (written) *written = total;
    return GOOFY_OK;
}

int goofy_container_write_file(const char* path, int kind, int format, unsigned int width, unsigned int height,
                               const void* blocks, size_t blocks_bytes)
{
    clear_error();
    int rc = check(kind, format, width, height, blocks_bytes);
    if (rc != GOOFY_OK) return rc;
    if (!path || (blocks_bytes && !blocks)) return fail(GOOFY_ERR_INVALID, "null argument");
    unsigned char hdr[kDdsHeader > kKtxHeader ? kDdsHeader : kKtxHeader];
    write_header(kind, format, width, height, blocks_bytes, hdr);
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(GOOFY_ERR_INVALID, std::string("cannot open ") + path);
    const bool ok = std::fwrite(hdr, 1, header_bytes(kind), f) == header_bytes(kind) &&
                    std::fwrite(blocks, 1, blocks_bytes, f) == blocks_bytes;
    if (std::fclose(f) != 0 || !ok) return fail(GOOFY_ERR_INVALID, std::string("write failed: ") + path);
    return GOOFY_OK;
}

}  // extern "C"
