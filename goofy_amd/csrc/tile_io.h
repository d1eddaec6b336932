// tile_io.h -- wave-uniform addressing of runs of consecutive 4x4 blocks in
// an RGBA8 image, for the encode, decode and round-trip quality kernels.
//
// A wave handles a run of `span` consecutive blocks (row-major block order)
// starting at block `first`.  Usual case (tile_near): one buffer resource per
// wave on the first pixel of the run's first block row, in scalar registers;
// a lane adds a 32-bit byte offset (block_offset) and pixel row r its
// wave-uniform r * pitch (the scalar offset of the buffer access), so no lane
// computes a 64-bit address.  With bw >= span a run covers at most two block
// rows, and a block of the second one sits 4*pitch + 16*bx < 5*pitch bytes
// (pitch >= 16*bw) past the first row's start; with pitch < 2^29 the lane
// offset plus 3 * pitch stays below 2^32.  Narrower images and larger pitches
// are for the callers' per-lane 64-bit path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace goofy_impl {

// Buffer resource over `base` (raw: no stride, no swizzle; dword 3 as for
// every gfx9 raw buffer).  Wave-uniform: it lives in four scalar registers.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_rsrc(const void* base)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, 0xFFFFFFFF, 0x00020000);
}

// Cache policy of a buffer access: 2 = nontemporal (streaming), the "nt" of
// global loads and stores.
template <bool NT>
constexpr int kBufAux = NT ? 2 : 0;

__device__ __forceinline__ bool tile_near(uint32_t bw, uint32_t span, uint64_t pitch)
{
    return bw >= span && (pitch >> 29) == 0;
}

// The run's first block row: a resource on its first pixel, the column of
// block `first`, and the byte step from (row, column bw) to (row + 1,
// column 0).  magic: (i * magic) >> 40 == i / bw for the image's blocks.
struct BlockRows {
    __amdgpu_buffer_rsrc_t rows;
    uint32_t bx0, wrap;
};

__device__ __forceinline__ BlockRows block_rows(const uint8_t* image, uint64_t pitch, uint32_t bw, uint64_t magic,
                                                uint32_t first)
{
    const uint32_t by0 = (uint32_t)(((uint64_t)first * magic) >> 40);
    BlockRows t;
    t.bx0 = first - by0 * bw;
    t.rows = raw_rsrc(image + (uint64_t)by0 * 4u * pitch);
    t.wrap = (uint32_t)(4u * pitch) - 16u * bw;
    return t;
}

// Byte offset of block first + k (k < span <= bw) from the resource's base.
__device__ __forceinline__ uint32_t block_offset(const BlockRows& t, uint32_t bw, uint32_t k)
{
    const uint32_t rel = t.bx0 + k;
    return rel * 16u + (rel >= bw ? t.wrap : 0u);
}

}  // namespace goofy_impl
