// ref_shim.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A thin C-ABI wrapper, compiled together with the reference encoder header
// GoofyTC/goofy_tc.h taken straight from the reference tree (include path set
// by oracle/Makefile; nothing of the reference is copied into this
// repository).  The result, oracle/_ref/libgoofy_ref.so, is used to pin the
// C restatement (oracle/goofy_oracle.c), to write the golden vectors under
// tests/golden/, and as the CPU baseline of the benchmark (the reference SSE2
// path, one horizontal stripe of block rows per host thread).  The
// reference's block decoders (Src/decoder.cpp, compiled from where it lies)
// are wrapped too, to pin the decoder oracle and the GPU decoders.
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "decoder.h"
#include "goofy_tc.h"

namespace {

using CompressFn = int (*)(unsigned char*, const unsigned char*, unsigned, unsigned, unsigned);

CompressFn pick(int fmt) { return fmt == 0 ? &goofy::compressDXT1 : &goofy::compressETC1; }

// The reference reads rows with aligned 16-byte loads (goofy_tc.h:177) and
// asserts 64-byte alignment in debug builds (:1072).  Inputs that do not meet
// that are staged through a 64-byte aligned copy here rather than crashing.
bool aligned_ok(const unsigned char* src, unsigned stride)
{
    return (reinterpret_cast<uintptr_t>(src) % 64) == 0 && (stride % 16) == 0;
}

}  // namespace

extern "C" {

int goofy_ref_compress(int fmt, unsigned char* dst, const unsigned char* src, unsigned width, unsigned height,
                       unsigned stride)
{
    CompressFn f = pick(fmt);
    if (width % 16 != 0 || height % 4 != 0 || aligned_ok(src, stride)) return f(dst, src, width, height, stride);
    const size_t pitch = (size_t)width * 4;
    void* tmp = nullptr;
    if (posix_memalign(&tmp, 64, pitch * height ? pitch * height : 64) != 0) return -3;
    for (unsigned y = 0; y < height; ++y) memcpy((unsigned char*)tmp + y * pitch, src + (size_t)y * stride, pitch);
    int rc = f(dst, (const unsigned char*)tmp, width, height, (unsigned)pitch);
    free(tmp);
    return rc;
}

// One stripe of block rows per thread: stripe t reads input + y0*stride and
// writes result + (y0/4)*(width/4)*8, exactly as a caller of the reference
// API would split an image across cores.
int goofy_ref_compress_mt(int fmt, unsigned char* dst, const unsigned char* src, unsigned width, unsigned height,
                          unsigned stride, int nthreads)
{
    if (width % 16 != 0) return -1;
    if (height % 4 != 0) return -2;
    if (!aligned_ok(src, stride)) return -4;
    CompressFn f = pick(fmt);
    const unsigned bh = height / 4;
    if (nthreads < 1) nthreads = 1;
    if ((unsigned)nthreads > bh) nthreads = bh ? (int)bh : 1;
    std::vector<std::thread> pool;
    pool.reserve(nthreads);
    for (int t = 0; t < nthreads; ++t) {
        const unsigned r0 = (unsigned)((uint64_t)bh * t / nthreads);
        const unsigned r1 = (unsigned)((uint64_t)bh * (t + 1) / nthreads);
        pool.emplace_back([=] {
            f(dst + (size_t)r0 * (width / 4) * 8, src + (size_t)r0 * 4 * stride, width, (r1 - r0) * 4, stride);
        });
    }
    for (auto& th : pool) th.join();
    return 0;
}

// Encode n independent blocks (each 64 bytes, 4 rows of 16 bytes) by laying
// them side by side in a 4-row image of width 4*ceil4(n) and calling the
// reference driver.  Blocks of one call are independent in the reference
// (goofySimdEncode only re-lays data across its 4 blocks), which the test
// suite checks by permuting block order.
int goofy_ref_encode_blocks(int fmt, const unsigned char* blocks, size_t n, unsigned char* out)
{
    const size_t npad = (n + 3) & ~(size_t)3;
    if (npad == 0) return 0;
    const size_t pitch = npad * 16;
    void* img = nullptr;
    if (posix_memalign(&img, 64, pitch * 4) != 0) return -3;
    memset(img, 0, pitch * 4);
    unsigned char* im = (unsigned char*)img;
    for (size_t b = 0; b < n; ++b)
        for (int r = 0; r < 4; ++r) memcpy(im + r * pitch + b * 16, blocks + b * 64 + r * 16, 16);
    std::vector<unsigned char> res(npad * 8);
    int rc = pick(fmt)(res.data(), im, (unsigned)(npad * 4), 4, (unsigned)pitch);
    memcpy(out, res.data(), n * 8);
    free(img);
    return rc;
}

// Decode one 8-byte block into 16 RGBA8 pixels (row-major, 64 bytes) with
// the reference harness's decoders (DecoderBC::decodeBlockDXT1 / ETC1,
// Src/decoder.cpp:929-971).
void goofy_ref_decode_block(int fmt, const unsigned char* block, unsigned char* out)
{
    if (fmt == 0) DecoderBC::decodeBlockDXT1(block, out, 16);
    else DecoderBC::decodeBlockETC1(block, out, 16);
}

}  // extern "C"
