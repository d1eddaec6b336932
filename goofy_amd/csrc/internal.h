// internal.h -- declarations shared by the host-side translation units of
// libgoofy_hip.so (not installed; the public C-ABI is include/goofy_hip.h).
#pragma once

#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

#include <string>

#include "goofy_hip.h"

namespace goofy_impl {

// Device-side texture descriptor of the flattened work-list (64 bytes).
// Texture t owns wave tiles [tile_begin, tile_begin + ceil(nblocks / tile)).
struct TexDesc {
    uint64_t src;      // device address of pixel (0,0)
    uint64_t dst;      // device address of block 0 (8 bytes per block, row-major)
    uint64_t pitch;    // bytes between pixel rows
    uint64_t magic;    // ceil(2^40 / bw): block row = (i * magic) >> 40
    uint32_t bw;       // blocks per row
    uint32_t nblocks;  // bw * bh
    uint32_t tile_begin;  // also in the work-list's prefix-sum array
    uint32_t fmt;         // GOOFY_FORMAT_DXT1 / GOOFY_FORMAT_ETC1
    uint32_t pad[4];
};
static_assert(sizeof(TexDesc) == 64, "TexDesc layout");

// Kernel launch configuration (what the autotuner searches).
struct LaunchCfg {
    int blocks_per_lane;  // 1, 2 or 4 blocks per lane (tile = 64 * bpl blocks per wave)
    int wg_waves;         // waves per workgroup: 1, 2, 4 or 8
    int waves_per_cu;     // persistent grid size (0 = one wave per tile, no persistence)
    int nontemporal;      // 1 = streaming (non-temporal) loads and stores
    int tile_order;       // 0 = linear, 1 = XCD-contiguous (kXcdCount groups of workgroups)
};

// The dispatcher hands workgroup b to XCD b % 8.  With tile_order 1, kernels
// renumber workgroups so that each XCD's workgroups take one contiguous
// eighth of the tiles, in order.
constexpr uint32_t kXcdCount = 8;

inline LaunchCfg to_cfg(const goofy_launch_config& u)
{
    return LaunchCfg{u.blocks_per_lane, u.wg_waves, u.waves_per_cu, u.nontemporal, u.tile_order};
}
inline goofy_launch_config to_public(const LaunchCfg& c)
{
    return goofy_launch_config{c.blocks_per_lane, c.wg_waves, c.waves_per_cu, c.nontemporal, c.tile_order};
}

// A small work-list passed by value in the kernel arguments (no copy to the
// device before the launch): up to kInlineTextures descriptors and their
// ntex+1 tile prefix sums.
constexpr int kInlineTextures = 16;
struct InlineWorklist {
    TexDesc d[kInlineTextures];
    uint32_t tile_begin[kInlineTextures + 1];
};

// A batch work-list in device memory: ntex descriptors followed by the
// ntex+1 prefix sums of their tile counts (uint32, last = total tiles).
inline size_t worklist_bytes(size_t ntex) { return ntex * sizeof(TexDesc) + (ntex + 1) * sizeof(uint32_t); }
inline uint32_t* worklist_tile_begin(TexDesc* descs, size_t ntex) { return reinterpret_cast<uint32_t*>(descs + ntex); }
inline const uint32_t* worklist_tile_begin(const TexDesc* descs, size_t ntex)
{
    return reinterpret_cast<const uint32_t*>(descs + ntex);
}

inline bool format_valid(int f) { return f >= GOOFY_FORMAT_DXT1 && f <= GOOFY_FORMAT_ETC2; }
// Bytes per block: 8 (DXT1, ETC1S) or 16 (DXT5, ETC2 RGBA8).
inline size_t block_bytes(int f) { return f >= GOOFY_FORMAT_DXT5 ? 16u : 8u; }

inline uint64_t div_magic(uint32_t bw) { return ((1ull << 40) + bw - 1) / bw; }

// True when (i * magic) >> 40 == i / bw for every i < nblocks.
inline bool magic_ok(uint64_t bw, uint64_t nblocks) { return bw > 0 && nblocks * bw < (1ull << 40) && nblocks / bw < (1ull << 24); }

// Most block rows one launch (or one work-list entry) may cover for an image
// bw blocks wide: block indices below 2^32 and an exact division magic
// (magic_ok).  Larger images are encoded as horizontal stripes of at most
// this many block rows; blocks are independent, so stripes change nothing in
// the output.  0 when a single block row is already beyond the range
// (bw >= 2^20, i.e. 4,194,304 pixels or more per row).
inline uint32_t max_stripe_brows(uint64_t bw)
{
    if (bw == 0 || bw >= (1ull << 20)) return 0;
    uint64_t r = ((1ull << 40) - 1) / (bw * bw);
    if (r > 0xFFFFFFFFull / bw) r = 0xFFFFFFFFull / bw;
    if (r > (1ull << 24) - 1) r = (1ull << 24) - 1;
    return (uint32_t)r;
}

// Kernel launchers (kernels.hip).
hipError_t launch_plain(int fmt, const LaunchCfg& cfg, const void* src, size_t pitch, void* dst, uint32_t bw,
                        uint32_t nblocks, hipStream_t stream, hipEvent_t start = nullptr, hipEvent_t stop = nullptr);
hipError_t launch_dual(uint32_t wg_waves, const void* src, size_t pitch, void* dst_dxt1, void* dst_etc1, uint32_t bw,
                       uint32_t nblocks, hipStream_t stream, hipEvent_t start = nullptr, hipEvent_t stop = nullptr);
hipError_t launch_dual_alpha(uint32_t wg_waves, const void* src, size_t pitch, void* dst_dxt5, void* dst_etc2,
                             uint32_t bw, uint32_t nblocks, hipStream_t stream, hipEvent_t start = nullptr,
                             hipEvent_t stop = nullptr);
// alpha: the work-list holds DXT5 / ETC2 RGBA8 textures too (the kernel
// instance with all four formats; without, DXT1 / ETC1S only).
hipError_t launch_batch(const LaunchCfg& cfg, const TexDesc* d_descs, uint32_t ntex, uint32_t ntiles, int num_cus,
                        hipStream_t stream, bool alpha = false);
hipError_t launch_batch_inline(const LaunchCfg& cfg, const InlineWorklist& wl, uint32_t ntex, uint32_t ntiles,
                               int num_cus, hipStream_t stream, bool alpha = false);
hipError_t launch_synth(const void* params, void* dst, uint32_t width, uint32_t height, size_t pitch,
                        hipStream_t stream);
hipError_t launch_decode(int fmt, const void* blocks, void* dst, uint32_t width, uint32_t height, size_t pitch,
                         hipStream_t stream);
// Quality sums: squared errors of bytes 0..3, and the sum of squared luma
// differences times 10^4 (an exact integer with the harness's weights).  The
// quality launchers take a device scratch buffer of quality_scratch_bytes():
// the final sums at offset 0, one partial per workgroup from
// kQualityPartialsOffset on.
struct QualitySums {
    unsigned long long e[4];
    unsigned long long y100;
};
constexpr size_t kQualityPartialsOffset = 64;
constexpr uint32_t kQualityMaxGroups = 1536;
inline uint32_t quality_grid(uint64_t units)
{
    const uint64_t g = (units + 255u) / 256u;
    return (uint32_t)(g < 1 ? 1 : (g > kQualityMaxGroups ? kQualityMaxGroups : g));
}
inline size_t quality_scratch_bytes(uint64_t units)
{
    return kQualityPartialsOffset + (size_t)quality_grid(units) * sizeof(QualitySums);
}
// A quality lane adds at most kQualityLaneUnits blocks (or 4-pixel quads) in
// 32-bit sums, so one launch covers at most this many units; the callers
// launch larger images as stripes of quality_stripe_rows(units per row) rows.
constexpr uint64_t kQualityLaneUnits = 2048;
constexpr uint64_t kQualityMaxLaunchUnits = kQualityLaneUnits * kQualityMaxGroups * 256u;
inline uint32_t quality_stripe_rows(uint64_t units_per_row)
{
    const uint64_t r = max_stripe_brows(units_per_row), cap = kQualityMaxLaunchUnits / (units_per_row ? units_per_row : 1);
    return (uint32_t)(r < cap ? r : cap);
}
// accumulate: add onto the sums already at offset 0 of `work` (later
// stripes of one image on the same stream) instead of replacing them.
// host_out: null, or pinned host memory for a copy of the sums at offset 0
// (one QualitySums), valid once the stream has drained.  Every partial the
// final kernel reads is written by the launch before it, and the first stripe
// replaces the sums, so `work` needs no clearing between calls.
hipError_t launch_sqerr(int fmt, const void* blocks, const void* src, size_t pitch, uint32_t width, uint32_t height,
                        void* work, bool accumulate, void* host_out, hipStream_t stream);
hipError_t launch_sqerr_images(const void* a, size_t pitch_a, const void* b, size_t pitch_b, uint32_t width,
                               uint32_t height, void* work, bool accumulate, void* host_out, hipStream_t stream);
// Size of the largest quality work area (kQualityMaxGroups partials), the
// size of the allocation-free call's fixed device array.
constexpr size_t kQualityNoAllocBytes = kQualityPartialsOffset + (size_t)kQualityMaxGroups * sizeof(QualitySums);
// Device address of `device`'s copy of that array (the current device of the
// calling thread), looked up once per device and kept.
hipError_t noalloc_scratch_address(int device, void** addr);
// getMsePsnr's figures (Src/main.cpp:403-469) from the integer sums.
void quality_figures(const QualitySums& sums, uint64_t pixels, goofy_quality* out);
hipError_t launch_rgb565_baseline(const void* src, size_t src_pitch, void* dst, size_t dst_pitch, uint32_t width,
                                  uint32_t height, hipStream_t stream, hipEvent_t start = nullptr,
                                  hipEvent_t stop = nullptr);

// Error reporting: every C-ABI function stores a message for its thread.
int fail(int code, const std::string& msg);
int fail_hip(hipError_t e, const char* what);
void clear_error();

// Device properties cached per device.
struct DeviceInfo {
    int device = -1;
    int num_cus = 0;
    size_t total_mem = 0;
    char arch[64] = {0};
};
int device_info(DeviceInfo* out);  // current device; returns GOOFY_OK or an error code

inline int tile_blocks(const LaunchCfg& c) { return 64 * c.blocks_per_lane; }
// Default launch shape for a launch of `blocks` 4x4 blocks on a device with
// num_cus compute units (used wherever the caller gives no shape).
LaunchCfg auto_cfg(uint64_t blocks, int num_cus);
bool cfg_valid(const LaunchCfg& c);

}  // namespace goofy_impl
