// goofy_tc.h -- drop-in replacement for the reference's public API.
//
// Same namespace, names, argument meaning, return codes and output layout as
// GoofyTC/goofy_tc.h:10-13 (implementation :1497-1557), but declarations
// only: the definitions live in libgoofy_hip.so and run the MI355X (gfx950)
// HIP kernels through the C-ABI of goofy_hip.h.  Unlike the reference
// header, including this file in several translation units is fine.
//
//   result  host buffer of (width/4)*(height/4)*8 bytes, blocks row-major and
//           tightly packed (no destination stride)
//   input   host RGBA8 pixels, `stride` bytes between rows (stride >= 4*width)
//   return  0 on success, -1 if width % 16 != 0, -2 if height % 4 != 0
//           (checked in that order, as in the reference).  Codes the
//           reference never returns: -3 HIP runtime error, -4 invalid
//           argument (null pointer or stride < 4*width), -5 device out of
//           memory, -6 no GPU.
#ifndef GOOFY_TC_DROPIN_H
#define GOOFY_TC_DROPIN_H

namespace goofy {
// Replaces goofy::compressDXT1, GoofyTC/goofy_tc.h:11 (definition :1497-1526).
int compressDXT1(unsigned char* result, const unsigned char* input, unsigned int width, unsigned int height,
                 unsigned int stride);
// Replaces goofy::compressETC1, GoofyTC/goofy_tc.h:12 (definition :1528-1557).
int compressETC1(unsigned char* result, const unsigned char* input, unsigned int width, unsigned int height,
                 unsigned int stride);

// No counterpart in the reference, whose README names a DXT5 / ETC2 alpha
// encoder as its next step: the same contract with 16-byte blocks, i.e.
// `result` holds (width/4)*(height/4)*16 bytes.  DXT5 is BC3 and ETC2 is ETC2
// RGBA8 (EAC alpha); each block is an alpha block followed by the block
// compressDXT1 / compressETC1 write for the same pixels.
int compressDXT5(unsigned char* result, const unsigned char* input, unsigned int width, unsigned int height,
                 unsigned int stride);
int compressETC2(unsigned char* result, const unsigned char* input, unsigned int width, unsigned int height,
                 unsigned int stride);
}  // namespace goofy

#endif
