// harness_shim.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A C-ABI wrapper around the reference benchmark harness, Src/main.cpp,
// which this translation unit includes straight from the reference tree
// (include path set by oracle/Makefile; nothing of the reference is copied
// into this repository).  The result, oracle/_ref/libgoofy_harness.so, pins
// the parts of this library that sit either side of the encoder against the
// harness's own code:
//   * PNG ingestion      -- loadPngAsRgba8 / destroyPng   (Src/main.cpp:258-349)
//   * DDS / KTX1 writers -- saveDds / saveKtx            (Src/main.cpp:154-220)
//   * image decoders     -- decompressDXT1 / DXT5 / ETC1 / ETC2 (Src/main.cpp:561-631)
//   * quality figures    -- getMsePsnr                   (Src/main.cpp:403-469)
//   * quality floor      -- generateDownsamlpedRgb565Test (Src/main.cpp:501-555)
//   * the plugin loop    -- runTestDXT1 / runTestETC1 over a CompressFunc_t
//                           (Src/main.cpp:644-743), and the harness's own
//                           goofy::compressDXT1 / compressETC1 to run in it
// The harness's main() comes along in the same object; nothing calls it.
#include "main.cpp"

#define GOOFY_HARNESS_API extern "C" __attribute__((visibility("default")))

// Decode a PNG the way the harness does.  Returns the 64-byte aligned RGBA8
// buffer (free it with goofy_harness_free) or null when lodepng fails or the
// size is refused (width % 16, height % 4).
GOOFY_HARNESS_API unsigned char* goofy_harness_load_png(const char* path, unsigned* width, unsigned* height)
{
    return loadPngAsRgba8(path, width, height);
}

GOOFY_HARNESS_API void goofy_harness_free(unsigned char* rgba8) { destroyPng(rgba8); }

// out: 14 doubles in MsePsnr field order -- mseR, mseG, mseB, mseA, mseMax,
// mseRGB, mseY, psnrR, psnrG, psnrB, psnrA, psnrMin, psnrRGB, psnrY.
GOOFY_HARNESS_API void goofy_harness_mse_psnr(const unsigned char* a, const unsigned char* b, unsigned width,
                                              unsigned height, double* out)
{
    const MsePsnr r = getMsePsnr(a, b, width, height);
    const double v[14] = {r.mseR,  r.mseG,  r.mseB,  r.mseA,  r.mseMax,  r.mseRGB, r.mseY,
                          r.psnrR, r.psnrG, r.psnrB, r.psnrA, r.psnrMin, r.psnrRGB, r.psnrY};
    for (int i = 0; i < 14; ++i) out[i] = v[i];
}

GOOFY_HARNESS_API void goofy_harness_rgb565_baseline(const unsigned char* in, unsigned width, unsigned height,
                                                     unsigned char* out)
{
    generateDownsamlpedRgb565Test(in, width, height, out);
}

// fmt 0 = DXT1, 1 = ETC1, 2 = DXT5, 3 = ETC2 RGBA8.  Whole image, tight rows,
// blocks row-major.
GOOFY_HARNESS_API void goofy_harness_decompress(int fmt, const unsigned char* blocks, unsigned width, unsigned height,
                                                unsigned char* rgba8)
{
    if (fmt == 0) decompressDXT1(blocks, width, height, rgba8);
    else if (fmt == 1) decompressETC1(blocks, width, height, rgba8);
    else if (fmt == 2) decompressDXT5(blocks, width, height, rgba8);
    else decompressETC2(blocks, width, height, rgba8);
}

// container 0 = DDS (FourCC DXT1, or DXT5 for fmt 2), 1 = KTX1 (fmt 0 = DXT1,
// 1 = ETC1, 2 = DXT5, 3 = ETC2 RGBA8).
GOOFY_HARNESS_API void goofy_harness_save(int container, int fmt, const char* path, const unsigned char* data,
                                          size_t size, unsigned width, unsigned height)
{
    static const uint32_t ktx[4] = {kKtxFormatDXT1, kKtxFormatETC1, kKtxFormatDXT5, kKtxFormatETC2};
    if (container == 0) saveDds(path, data, size, width, height, fmt == 2 ? kDdsFormatDXT5 : kDdsFormatDXT1);
    else saveKtx(path, data, size, width, height, ktx[fmt & 3]);
}

// The harness's own per-encoder loop, runTestDXT1 / runTestETC1
// (Src/main.cpp:646-743), around any CompressFunc_t (:644): `iterations`
// timed calls with the best kept, then the harness's decode, getMsePsnr and
// its .tga and .dds / .ktx files under ./test-results/ of the working
// directory.  fmt 0 = DXT1, 1 = ETC1.  time_us: the best call; mse_psnr: 14
// doubles in the order of goofy_harness_mse_psnr; dst: the last call's blocks.
GOOFY_HARNESS_API void goofy_harness_run_test(int fmt, const char* encoder, const char* image, CompressFunc_t func,
                                              unsigned iterations, unsigned char* src, unsigned width,
                                              unsigned height, unsigned stride, unsigned char* dst, size_t dst_size,
                                              double* time_us, double* mse_psnr)
{
    Timer timer;
    std::vector<unsigned char> scratch((size_t)width * height * 4);
    const TestResult r =
        fmt == 0 ? runTestDXT1(encoder, image, func, timer, iterations, dst, dst_size, src, width, height, stride,
                               scratch.data())
                 : runTestETC1(encoder, image, func, timer, iterations, dst, dst_size, src, width, height, stride,
                               scratch.data());
    *time_us = r.timeInMicroSeconds;
    const MsePsnr& q = r.msePsnr;
    const double v[14] = {q.mseR,  q.mseG,  q.mseB,  q.mseA,  q.mseMax,  q.mseRGB, q.mseY,
                          q.psnrR, q.psnrG, q.psnrB, q.psnrA, q.psnrMin, q.psnrRGB, q.psnrY};
    for (int i = 0; i < 14; ++i) mse_psnr[i] = v[i];
}

// goofy::compressDXT1 / compressETC1 as the harness compiles them (its
// "simd_goofy" encoder, Src/main.cpp:925-929).
GOOFY_HARNESS_API CompressFunc_t goofy_harness_goofy_encoder(int fmt)
{
    return fmt == 0 ? goofy::compressDXT1 : goofy::compressETC1;
}
