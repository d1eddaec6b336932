// probe.hip -- roofline calibration: the encoder's memory traffic without its
// arithmetic.  Lane i reads the four 16-byte rows of block i exactly as
// k_encode_plain does with one block per lane (same tiles, same streaming
// hints, workgroups of 1-8 waves) and writes 8 bytes (an XOR fold of the 64),
// so its best duration over the workgroup sizes is the floor the encode
// kernels can reach on a given device; bench.py reports the encoder against it.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "internal.h"

namespace goofy_impl {

namespace {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

__global__ __launch_bounds__(512) void k_traffic_probe(const uint8_t* __restrict__ src, uint64_t pitch,
                                                       uint8_t* __restrict__ dst, uint32_t bw, uint32_t nblocks,
                                                       uint64_t magic)
{
    const uint32_t wave = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    const uint32_t i = wave * 64u + (threadIdx.x & 63u);
    if (i >= nblocks) return;
    const uint32_t by = (uint32_t)(((uint64_t)i * magic) >> 40), bx = i - by * bw;
    const uint8_t* b = src + (uint64_t)by * 4u * pitch + (uint64_t)bx * 16u;
    u32x4 acc = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(b));
#pragma unroll
    for (int r = 1; r < 4; ++r) acc ^= __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(b + r * pitch));
    const u32x2 o = {acc.x ^ acc.y, acc.z ^ acc.w};
    __builtin_nontemporal_store(o, reinterpret_cast<u32x2*>(dst + (uint64_t)i * 8u));
}

}  // namespace

}  // namespace goofy_impl

using namespace goofy_impl;

extern "C" int goofy_hip_traffic_probe(void* d_out, const void* d_input, unsigned int width, unsigned int height,
                                       size_t pitch, int wg_waves, void* stream, void* start_event, void* stop_event)
{
    clear_error();
    if (wg_waves == 0) wg_waves = 8;
    if (wg_waves != 1 && wg_waves != 2 && wg_waves != 4 && wg_waves != 8)
        return fail(GOOFY_ERR_INVALID, "wg_waves must be 1, 2, 4 or 8");
    if (width % 4 != 0) return fail(GOOFY_ERR_WIDTH, "width must be a multiple of 4");
    if (height % 4 != 0) return fail(GOOFY_ERR_HEIGHT, "height must be a multiple of 4");
    if (width == 0 || height == 0) return GOOFY_OK;
    if (!d_out || !d_input || pitch < (size_t)width * 4 || pitch % 16 != 0 || (uintptr_t)d_input % 16 != 0 ||
        (uintptr_t)d_out % 8 != 0)
        return fail(GOOFY_ERR_INVALID, "bad buffers or pitch");
    const uint32_t bw = width / 4, nblocks = bw * (height / 4);
    if (!magic_ok(bw, nblocks)) return fail(GOOFY_ERR_TOO_LARGE, "image too large");
    const uint32_t threads = 64u * (uint32_t)wg_waves;
    const dim3 grid((nblocks + threads - 1u) / threads), block(threads);
    hipExtLaunchKernelGGL(k_traffic_probe, grid, block, 0, (hipStream_t)stream, (hipEvent_t)start_event,
                          (hipEvent_t)stop_event, 0, (const uint8_t*)d_input, (uint64_t)pitch, (uint8_t*)d_out, bw,
                          nblocks, div_magic(bw));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail_hip(e, "probe kernel launch");
    return GOOFY_OK;
}
