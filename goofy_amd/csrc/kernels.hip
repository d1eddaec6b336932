// kernels.hip -- gfx950 kernels of libgoofy_hip.so and their launchers.
//
// Work decomposition: one lane encodes one 4x4 block (blocks_per_lane of
// them with the autotuned shape); one wavefront owns a "tile" of 64*bpl
// consecutive blocks of one texture in row-major block order.  Lane l of a
// wave reads rows 0..3 of block (tile + l) as four 16-byte loads, so each
// load instruction of a wave covers 1 KiB of contiguous pixel row (64 blocks
// x 16 B) -- fully coalesced without an LDS staging pass -- and the 8-byte
// results of a wave land in 512 contiguous bytes.
//
// This replaces the reference driver loop (GoofyTC/goofy_tc.h:1514-1524,
// 4 blocks per call) and goofySimdEncode (:1069-1494); the arithmetic of a
// block is in block_codec.h.
#include <algorithm>

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "block_codec.h"
#include "internal.h"
#include "synth.h"
#include "tile_io.h"

namespace goofy_impl {

using goofy_dev::BlockClass;
using goofy_dev::classify;
using goofy_dev::encode_block;
using goofy_dev::encode_block16;
using goofy_dev::pack_dxt1;
using goofy_dev::pack_etc1;
using goofy_dev::AlphaRows;
using goofy_dev::alpha_rows;
using goofy_dev::bc3_alpha;
using goofy_dev::eac_alpha;

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

template <bool NT>
__device__ __forceinline__ u32x4 load16(const uint8_t* p)
{
    if constexpr (NT) return __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p));
    else return *reinterpret_cast<const u32x4*>(p);
}

template <bool NT>
__device__ __forceinline__ void store8(uint8_t* p, uint2 v)
{
    const u32x2 w = {v.x, v.y};
    if constexpr (NT) __builtin_nontemporal_store(w, reinterpret_cast<u32x2*>(p));
    else *reinterpret_cast<u32x2*>(p) = w;
}

template <bool NT>
__device__ __forceinline__ void store16(uint8_t* p, uint4 v)
{
    const u32x4 w = {v.x, v.y, v.z, v.w};
    if constexpr (NT) __builtin_nontemporal_store(w, reinterpret_cast<u32x4*>(p));
    else *reinterpret_cast<u32x4*>(p) = w;
}

// Load the four pixel rows of the blocks first + 64k + lane (k < BPL) of a
// texture; lanes past nblocks get zeros.  Addressing as in tile_io.h, over
// runs of 64*BPL blocks; narrower textures and pitches of 2^29 and more take
// a per-lane 64-bit address.
template <int BPL, bool NT>
__device__ __forceinline__ void load_tile(const uint8_t* __restrict__ src, uint64_t pitch, uint32_t bw,
                                          uint64_t magic, uint32_t nblocks, uint32_t first, uint32_t lane,
                                          u32x4 (&px)[BPL][4])
{
    if (tile_near(bw, 64u * BPL, pitch)) {
        const BlockRows t = block_rows(src, pitch, bw, magic, first);
#pragma unroll
        for (int k = 0; k < BPL; ++k) {
            const uint32_t off = block_offset(t, bw, 64u * k + lane);
            if (first + 64u * k + lane < nblocks) {
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    px[k][r] = __builtin_amdgcn_raw_buffer_load_b128(t.rows, off, (uint32_t)pitch * r, kBufAux<NT>);
            } else {
#pragma unroll
                for (int r = 0; r < 4; ++r) px[k][r] = u32x4{0u, 0u, 0u, 0u};
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < BPL; ++k) {
            const uint32_t i = first + 64u * k + lane;
            if (i < nblocks) {
                const uint32_t by = (uint32_t)(((uint64_t)i * magic) >> 40);
                const uint32_t bx = i - by * bw;
                const uint8_t* b = src + (uint64_t)by * 4u * pitch + (uint64_t)bx * 16u;
#pragma unroll
                for (int r = 0; r < 4; ++r) px[k][r] = load16<NT>(b + r * pitch);
            } else {
#pragma unroll
                for (int r = 0; r < 4; ++r) px[k][r] = u32x4{0u, 0u, 0u, 0u};
            }
        }
    }
}

// Results of a tile: a buffer resource on the tile's first output block, a
// lane offset below 64*BPL*16 bytes.
template <bool NT>
__device__ __forceinline__ void store8_at(__amdgpu_buffer_rsrc_t out, uint32_t off, uint2 v)
{
    const u32x2 w = {v.x, v.y};
    __builtin_amdgcn_raw_buffer_store_b64(w, out, off, 0, kBufAux<NT>);
}

template <bool NT>
__device__ __forceinline__ void store16_at(__amdgpu_buffer_rsrc_t out, uint32_t off, uint4 v)
{
    const u32x4 w = {v.x, v.y, v.z, v.w};
    __builtin_amdgcn_raw_buffer_store_b128(w, out, off, 0, kBufAux<NT>);
}

__device__ __forceinline__ void block_words(const u32x4 (&v)[4], uint32_t (&p)[16])
{
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        p[4 * r] = v[r].x;
        p[4 * r + 1] = v[r].y;
        p[4 * r + 2] = v[r].z;
        p[4 * r + 3] = v[r].w;
    }
}

// Encode the tile of 64*BPL blocks starting at block `first` of a texture.
// DXT1 / ETC1S write 8 bytes per block, DXT5 / ETC2 RGBA8 16 (a wave's
// results then fill 1 KiB contiguously), at a wave-uniform base plus a
// 32-bit lane offset.
// All loads of the tile are issued before any arithmetic so that a wave has
// 64*BPL*64 bytes in flight.
template <int FMT, int BPL, bool NT>
__device__ __forceinline__ void encode_tile(const uint8_t* __restrict__ src, uint64_t pitch, uint8_t* __restrict__ dst,
                                            uint32_t bw, uint64_t magic, uint32_t nblocks, uint32_t first,
                                            uint32_t lane)
{
    u32x4 px[BPL][4];
    load_tile<BPL, NT>(src, pitch, bw, magic, nblocks, first, lane, px);
    constexpr uint32_t kOut = FMT >= goofy_dev::kDXT5 ? 16u : 8u;
    const __amdgpu_buffer_rsrc_t out = raw_rsrc(dst + (uint64_t)first * kOut);
#pragma unroll
    for (int k = 0; k < BPL; ++k) {
        if (first + 64u * k + lane < nblocks) {
            uint32_t p[16];
            block_words(px[k], p);
            const uint32_t o = (64u * k + lane) * kOut;
            if constexpr (FMT >= goofy_dev::kDXT5) store16_at<NT>(out, o, encode_block16<FMT>(p));
            else store8_at<NT>(out, o, encode_block<FMT>(p));
        }
    }
}

// Index of this wave in the launch.  With order 1 the workgroups are
// renumbered so that those the dispatcher places on XCD x (b % 8 == x) take
// the x-th contiguous share of the launch, in order: each XCD then streams
// through one contiguous region instead of every eighth workgroup's tiles.
// A bijection on [0, gridDim.x): r = n % 8 XCDs hold q + 1 workgroups, the
// rest q.  All operands are wave-uniform (scalar unit).
// wpg = waves per workgroup.  The batch kernels take it as an argument; the
// others read blockDim.x, a 16-bit field of the dispatch packet, which this
// target reads with a vector load at the head of every wave.  Measured both
// ways: on the DXT1 / ETC1S kernels at 8192^2 (HBM-bound) the read is 2%
// faster than the argument, and no different on the alpha ones; on the
// 4096-texture batch, whose waves first search the work-list, the argument is
// 2-3% faster.
__device__ __forceinline__ uint32_t wave_index(uint32_t order, uint32_t wpg)
{
    uint32_t b = blockIdx.x;
    if (order) {
        const uint32_t n = gridDim.x, q = n / kXcdCount, r = n % kXcdCount;
        const uint32_t x = b % kXcdCount, k = b / kXcdCount;
        b = x < r ? x * (q + 1u) + k : r * (q + 1u) + (x - r) * q + k;
    }
    return __builtin_amdgcn_readfirstlane(b * wpg + (threadIdx.x >> 6));
}

__device__ __forceinline__ uint32_t wave_index(uint32_t order) { return wave_index(order, blockDim.x >> 6); }

// One image; wave w encodes tile w.
template <int FMT, int BPL, bool NT>
__global__ __launch_bounds__(512) void k_encode_plain(const uint8_t* __restrict__ src, uint64_t pitch,
                                                      uint8_t* __restrict__ dst, uint32_t bw, uint32_t nblocks,
                                                      uint64_t magic, uint32_t order)
{
    const uint32_t wave = wave_index(order);
    encode_tile<FMT, BPL, NT>(src, pitch, dst, bw, magic, nblocks, wave * 64u * BPL, threadIdx.x & 63u);
}

// One image into both formats: each block is loaded and classified once
// (bounding box, luma, range, thresholds, per-pixel masks) and packed as DXT1
// and as ETC1S.  72 -> 80 bytes per block instead of 2 x 72 for two plain
// launches.  One block per lane, streaming loads and stores.
__global__ __launch_bounds__(512) void k_encode_dual(const uint8_t* __restrict__ src, uint64_t pitch,
                                                     uint8_t* __restrict__ dst_dxt1, uint8_t* __restrict__ dst_etc1,
                                                     uint32_t bw, uint32_t nblocks, uint64_t magic)
{
    const uint32_t first = wave_index(0u) * 64u, lane = threadIdx.x & 63u;
    u32x4 v[1][4];
    load_tile<1, true>(src, pitch, bw, magic, nblocks, first, lane, v);
    if (first + lane >= nblocks) return;
    uint32_t p[16];
    block_words(v[0], p);
    const BlockClass c = classify(p);
    store8_at<true>(raw_rsrc(dst_dxt1 + (uint64_t)first * 8u), lane * 8u, pack_dxt1(c));
    store8_at<true>(raw_rsrc(dst_etc1 + (uint64_t)first * 8u), lane * 8u, pack_etc1(p, c));
}

// One image into DXT5 and ETC2 RGBA8: each block is loaded once, and the
// colour classification and the alpha rows are computed once for both.
// 64 + 16 + 16 = 96 bytes per block instead of 2 x 80 for two plain launches.
__global__ __launch_bounds__(512) void k_encode_dual_alpha(const uint8_t* __restrict__ src, uint64_t pitch,
                                                           uint8_t* __restrict__ dst_dxt5,
                                                           uint8_t* __restrict__ dst_etc2, uint32_t bw,
                                                           uint32_t nblocks, uint64_t magic)
{
    const uint32_t first = wave_index(0u) * 64u, lane = threadIdx.x & 63u;
    u32x4 v[1][4];
    load_tile<1, true>(src, pitch, bw, magic, nblocks, first, lane, v);
    if (first + lane >= nblocks) return;
    uint32_t p[16];
    block_words(v[0], p);
    uint32_t amn, amx;
    const BlockClass c = classify(p, goofy_dev::block_stats_alpha(p, amn, amx));
    const AlphaRows a = alpha_rows(p, amn, amx);
    const uint2 d = pack_dxt1(c), ba = bc3_alpha(a);
    store16_at<true>(raw_rsrc(dst_dxt5 + (uint64_t)first * 16u), lane * 16u, make_uint4(ba.x, ba.y, d.x, d.y));
    const uint2 e = pack_etc1(p, c), ea = eac_alpha(a);
    store16_at<true>(raw_rsrc(dst_etc2 + (uint64_t)first * 16u), lane * 16u, make_uint4(ea.x, ea.y, e.x, e.y));
}

// Many textures.  tile_begin[0..ntex] is the prefix sum of the textures'
// tile counts (tile_begin[ntex] = ntiles); texture d owns tiles
// [tile_begin[d], tile_begin[d+1]).  Wave w encodes the contiguous tile range
// [w*tpw, (w+1)*tpw): one wave-uniform binary search over the compact
// prefix-sum array (scalar loads) finds the first tile's texture, after which
// the wave steps forward across texture boundaries.  tpw = 1 is one wave per
// tile; tpw > 1 is a persistent grid.  Formats are uniform per tile, so the
// format branch never diverges.
// ALPHA: the work-list may hold DXT5 / ETC2 RGBA8 textures as well; batches
// of DXT1 / ETC1S textures only run the instance without those branches.
template <int BPL, bool NT, bool ALPHA, class Descs, class Begins>
__device__ __forceinline__ void batch_body(const Descs& descs, const Begins& tile_begin, uint32_t ntex,
                                           uint32_t ntiles, uint32_t tpw, uint32_t order, uint32_t wpg)
{
    const uint32_t wave = wave_index(order, wpg);
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t t = wave * tpw;
    const uint32_t t_end = min(t + tpw, ntiles);
    if (t >= t_end) return;
    uint32_t lo = 0, hi = ntex;
    while (hi - lo > 1u) {
        const uint32_t m = (lo + hi) >> 1;
        if (tile_begin[m] <= t) lo = m;
        else hi = m;
    }
    uint32_t d = lo, begin = tile_begin[d], next = tile_begin[d + 1];
    for (; t < t_end; ++t) {
        while (t >= next) {  // every texture owns at least one tile, and next <= ntiles
            ++d;
            begin = next;
            next = tile_begin[d + 1];
        }
        const TexDesc& D = descs[d];
        const uint32_t first = (t - begin) * 64u * BPL;
        const uint8_t* src = reinterpret_cast<const uint8_t*>(D.src);
        uint8_t* dst = reinterpret_cast<uint8_t*>(D.dst);
        if constexpr (ALPHA) {
            switch (D.fmt) {
            case 0: encode_tile<0, BPL, NT>(src, D.pitch, dst, D.bw, D.magic, D.nblocks, first, lane); break;
            case 1: encode_tile<1, BPL, NT>(src, D.pitch, dst, D.bw, D.magic, D.nblocks, first, lane); break;
            case 2: encode_tile<2, BPL, NT>(src, D.pitch, dst, D.bw, D.magic, D.nblocks, first, lane); break;
            default: encode_tile<3, BPL, NT>(src, D.pitch, dst, D.bw, D.magic, D.nblocks, first, lane);
            }
        } else {
            if (D.fmt == 0) encode_tile<0, BPL, NT>(src, D.pitch, dst, D.bw, D.magic, D.nblocks, first, lane);
            else encode_tile<1, BPL, NT>(src, D.pitch, dst, D.bw, D.magic, D.nblocks, first, lane);
        }
    }
}

template <int BPL, bool NT, bool ALPHA>
__global__ __launch_bounds__(512) void k_encode_batch(const TexDesc* __restrict__ descs,
                                                      const uint32_t* __restrict__ tile_begin, uint32_t ntex,
                                                      uint32_t ntiles, uint32_t tpw, uint32_t order, uint32_t wpg)
{
    batch_body<BPL, NT, ALPHA>(descs, tile_begin, ntex, ntiles, tpw, order, wpg);
}

// The same with the work-list in the kernel arguments (small groups).
template <int BPL, bool NT, bool ALPHA>
__global__ __launch_bounds__(512) void k_encode_batch_inline(const InlineWorklist wl, uint32_t ntex, uint32_t ntiles,
                                                             uint32_t tpw, uint32_t order, uint32_t wpg)
{
    batch_body<BPL, NT, ALPHA>(wl.d, wl.tile_begin, ntex, ntiles, tpw, order, wpg);
}

__global__ __launch_bounds__(256) void k_synth(goofy_synth::Params P, uint8_t* __restrict__ dst, uint32_t width,
                                               uint32_t height, uint64_t pitch)
{
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t y = blockIdx.y;
    if (x >= width || y >= height) return;
    *reinterpret_cast<uint32_t*>(dst + (uint64_t)y * pitch + (uint64_t)x * 4u) = goofy_synth::pixel(P, x, y);
}

// ---- launchers ---------------------------------------------------------------

// With a start/stop event pair the launch goes through hipExtLaunchKernelGGL,
// which has the dispatch packet itself record both events: the interval is
// the kernel's execution, without marker packets around it.
template <int FMT, int BPL, bool NT>
static hipError_t plain_go(dim3 grid, dim3 block, const void* src, size_t pitch, void* dst, uint32_t bw,
                           uint32_t nblocks, uint64_t magic, uint32_t order, hipStream_t s, hipEvent_t e0,
                           hipEvent_t e1)
{
    if (e0 || e1)
        hipExtLaunchKernelGGL((k_encode_plain<FMT, BPL, NT>), grid, block, 0, s, e0, e1, 0, (const uint8_t*)src,
                              (uint64_t)pitch, (uint8_t*)dst, bw, nblocks, magic, order);
    else
        hipLaunchKernelGGL((k_encode_plain<FMT, BPL, NT>), grid, block, 0, s, (const uint8_t*)src, (uint64_t)pitch,
                           (uint8_t*)dst, bw, nblocks, magic, order);
    return hipGetLastError();
}

template <int FMT>
static hipError_t plain_bpl(const LaunchCfg& c, dim3 grid, dim3 block, const void* src, size_t pitch, void* dst,
                            uint32_t bw, uint32_t nblocks, uint64_t magic, hipStream_t s, hipEvent_t e0,
                            hipEvent_t e1)
{
    const bool nt = c.nontemporal != 0;
    const uint32_t o = (uint32_t)c.tile_order;
    switch (c.blocks_per_lane) {
    case 1:
        return nt ? plain_go<FMT, 1, true>(grid, block, src, pitch, dst, bw, nblocks, magic, o, s, e0, e1)
                  : plain_go<FMT, 1, false>(grid, block, src, pitch, dst, bw, nblocks, magic, o, s, e0, e1);
    case 2:
        return nt ? plain_go<FMT, 2, true>(grid, block, src, pitch, dst, bw, nblocks, magic, o, s, e0, e1)
                  : plain_go<FMT, 2, false>(grid, block, src, pitch, dst, bw, nblocks, magic, o, s, e0, e1);
    default:
        return nt ? plain_go<FMT, 4, true>(grid, block, src, pitch, dst, bw, nblocks, magic, o, s, e0, e1)
                  : plain_go<FMT, 4, false>(grid, block, src, pitch, dst, bw, nblocks, magic, o, s, e0, e1);
    }
}

hipError_t launch_plain(int fmt, const LaunchCfg& c, const void* src, size_t pitch, void* dst, uint32_t bw,
                        uint32_t nblocks, hipStream_t stream, hipEvent_t start, hipEvent_t stop)
{
    if (nblocks == 0) return hipSuccess;
    const uint32_t tile = 64u * (uint32_t)c.blocks_per_lane;
    const uint32_t ntiles = (nblocks + tile - 1) / tile;
    const uint32_t wg = (uint32_t)c.wg_waves;
    const dim3 grid((ntiles + wg - 1) / wg), block(64u * wg);
    const uint64_t magic = div_magic(bw);
    switch (fmt) {
    case GOOFY_FORMAT_DXT1: return plain_bpl<0>(c, grid, block, src, pitch, dst, bw, nblocks, magic, stream, start, stop);
    case GOOFY_FORMAT_ETC1: return plain_bpl<1>(c, grid, block, src, pitch, dst, bw, nblocks, magic, stream, start, stop);
    case GOOFY_FORMAT_DXT5: return plain_bpl<2>(c, grid, block, src, pitch, dst, bw, nblocks, magic, stream, start, stop);
    default: return plain_bpl<3>(c, grid, block, src, pitch, dst, bw, nblocks, magic, stream, start, stop);
    }
}

hipError_t launch_dual(uint32_t wg_waves, const void* src, size_t pitch, void* dst_dxt1, void* dst_etc1, uint32_t bw,
                       uint32_t nblocks, hipStream_t stream, hipEvent_t start, hipEvent_t stop)
{
    if (nblocks == 0) return hipSuccess;
    const uint32_t ntiles = (nblocks + 63u) / 64u;
    const dim3 grid((ntiles + wg_waves - 1u) / wg_waves), block(64u * wg_waves);
    const uint64_t magic = div_magic(bw);
    if (start || stop)
        hipExtLaunchKernelGGL(k_encode_dual, grid, block, 0, stream, start, stop, 0, (const uint8_t*)src,
                              (uint64_t)pitch, (uint8_t*)dst_dxt1, (uint8_t*)dst_etc1, bw, nblocks, magic);
    else
        hipLaunchKernelGGL(k_encode_dual, grid, block, 0, stream, (const uint8_t*)src, (uint64_t)pitch,
                           (uint8_t*)dst_dxt1, (uint8_t*)dst_etc1, bw, nblocks, magic);
    return hipGetLastError();
}

hipError_t launch_dual_alpha(uint32_t wg_waves, const void* src, size_t pitch, void* dst_dxt5, void* dst_etc2,
                             uint32_t bw, uint32_t nblocks, hipStream_t stream, hipEvent_t start, hipEvent_t stop)
{
    if (nblocks == 0) return hipSuccess;
    const uint32_t ntiles = (nblocks + 63u) / 64u;
    const dim3 grid((ntiles + wg_waves - 1u) / wg_waves), block(64u * wg_waves);
    const uint64_t magic = div_magic(bw);
    if (start || stop)
        hipExtLaunchKernelGGL(k_encode_dual_alpha, grid, block, 0, stream, start, stop, 0, (const uint8_t*)src,
                              (uint64_t)pitch, (uint8_t*)dst_dxt5, (uint8_t*)dst_etc2, bw, nblocks, magic);
    else
        hipLaunchKernelGGL(k_encode_dual_alpha, grid, block, 0, stream, (const uint8_t*)src, (uint64_t)pitch,
                           (uint8_t*)dst_dxt5, (uint8_t*)dst_etc2, bw, nblocks, magic);
    return hipGetLastError();
}

template <int BPL, bool ALPHA>
static hipError_t batch_go(bool nt, dim3 grid, dim3 block, const TexDesc* d, const uint32_t* tb, uint32_t ntex,
                           uint32_t ntiles, uint32_t tpw, uint32_t order, hipStream_t s)
{
    const uint32_t wpg = block.x / 64u;
    if (nt)
        hipLaunchKernelGGL((k_encode_batch<BPL, true, ALPHA>), grid, block, 0, s, d, tb, ntex, ntiles, tpw, order, wpg);
    else
        hipLaunchKernelGGL((k_encode_batch<BPL, false, ALPHA>), grid, block, 0, s, d, tb, ntex, ntiles, tpw, order, wpg);
    return hipGetLastError();
}

template <int BPL>
static hipError_t batch_nt(bool nt, bool alpha, dim3 grid, dim3 block, const TexDesc* d, const uint32_t* tb,
                           uint32_t ntex, uint32_t ntiles, uint32_t tpw, uint32_t order, hipStream_t s)
{
    return alpha ? batch_go<BPL, true>(nt, grid, block, d, tb, ntex, ntiles, tpw, order, s)
                 : batch_go<BPL, false>(nt, grid, block, d, tb, ntex, ntiles, tpw, order, s);
}

hipError_t launch_batch(const LaunchCfg& c, const TexDesc* d_descs, uint32_t ntex, uint32_t ntiles, int num_cus,
                        hipStream_t stream, bool alpha)
{
    if (ntiles == 0 || ntex == 0) return hipSuccess;
    const uint32_t* tb = worklist_tile_begin(d_descs, ntex);
    uint64_t waves = ntiles;
    if (c.waves_per_cu > 0) waves = std::min<uint64_t>(waves, (uint64_t)num_cus * (uint64_t)c.waves_per_cu);
    const uint32_t tpw = (uint32_t)((ntiles + waves - 1) / waves);
    waves = (ntiles + tpw - 1) / tpw;
    const uint32_t wg = (uint32_t)c.wg_waves;
    const dim3 grid((uint32_t)((waves + wg - 1) / wg)), block(64u * wg);
    switch (c.blocks_per_lane) {
    case 1: return batch_nt<1>(c.nontemporal, alpha, grid, block, d_descs, tb, ntex, ntiles, tpw, c.tile_order, stream);
    case 2: return batch_nt<2>(c.nontemporal, alpha, grid, block, d_descs, tb, ntex, ntiles, tpw, c.tile_order, stream);
    default: return batch_nt<4>(c.nontemporal, alpha, grid, block, d_descs, tb, ntex, ntiles, tpw, c.tile_order, stream);
    }
}

template <int BPL, bool ALPHA>
static hipError_t batch_inline_go(bool nt, dim3 grid, dim3 block, const InlineWorklist& wl, uint32_t ntex,
                                  uint32_t ntiles, uint32_t tpw, uint32_t order, hipStream_t s)
{
    const uint32_t wpg = block.x / 64u;
    if (nt)
        hipLaunchKernelGGL((k_encode_batch_inline<BPL, true, ALPHA>), grid, block, 0, s, wl, ntex, ntiles, tpw, order,
                           wpg);
    else
        hipLaunchKernelGGL((k_encode_batch_inline<BPL, false, ALPHA>), grid, block, 0, s, wl, ntex, ntiles, tpw, order,
                           wpg);
    return hipGetLastError();
}

template <int BPL>
static hipError_t batch_inline_nt(bool nt, bool alpha, dim3 grid, dim3 block, const InlineWorklist& wl, uint32_t ntex,
                                  uint32_t ntiles, uint32_t tpw, uint32_t order, hipStream_t s)
{
    return alpha ? batch_inline_go<BPL, true>(nt, grid, block, wl, ntex, ntiles, tpw, order, s)
                 : batch_inline_go<BPL, false>(nt, grid, block, wl, ntex, ntiles, tpw, order, s);
}

hipError_t launch_batch_inline(const LaunchCfg& c, const InlineWorklist& wl, uint32_t ntex, uint32_t ntiles,
                               int num_cus, hipStream_t stream, bool alpha)
{
    if (ntiles == 0 || ntex == 0) return hipSuccess;
    if (ntex > (uint32_t)kInlineTextures) return hipErrorInvalidValue;
    uint64_t waves = ntiles;
    if (c.waves_per_cu > 0) waves = std::min<uint64_t>(waves, (uint64_t)num_cus * (uint64_t)c.waves_per_cu);
    const uint32_t tpw = (uint32_t)((ntiles + waves - 1) / waves);
    waves = (ntiles + tpw - 1) / tpw;
    const uint32_t wg = (uint32_t)c.wg_waves;
    const dim3 grid((uint32_t)((waves + wg - 1) / wg)), block(64u * wg);
    switch (c.blocks_per_lane) {
    case 1: return batch_inline_nt<1>(c.nontemporal, alpha, grid, block, wl, ntex, ntiles, tpw, c.tile_order, stream);
    case 2: return batch_inline_nt<2>(c.nontemporal, alpha, grid, block, wl, ntex, ntiles, tpw, c.tile_order, stream);
    default: return batch_inline_nt<4>(c.nontemporal, alpha, grid, block, wl, ntex, ntiles, tpw, c.tile_order, stream);
    }
}

hipError_t launch_synth(const void* params, void* dst, uint32_t width, uint32_t height, size_t pitch,
                        hipStream_t stream)
{
    if (width == 0 || height == 0) return hipSuccess;
    const goofy_synth::Params P = *static_cast<const goofy_synth::Params*>(params);
    const dim3 grid((width + 255u) / 256u, height), block(256);
    hipLaunchKernelGGL(k_synth, grid, block, 0, stream, P, (uint8_t*)dst, width, height, (uint64_t)pitch);
    return hipGetLastError();
}

}  // namespace goofy_impl
