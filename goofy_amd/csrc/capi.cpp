// capi.cpp -- single-image entry points of libgoofy_hip.so: the drop-in
// goofy::compressDXT1 / compressETC1 (host memory, synchronous), the device
// pointer / stream variant, error reporting and the synthetic-input fill.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "goofy_hip.h"
#include "goofy_tc.h"
#include "copy_pool.h"
#include "internal.h"
#include "synth.h"

namespace goofy_impl {

// ---- errors -----------------------------------------------------------------

static thread_local std::string t_last_error;

int fail(int code, const std::string& msg)
{
    t_last_error = msg;
    return code;
}

int fail_hip(hipError_t e, const char* what)
{
    const int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? GOOFY_ERR_OUT_OF_MEMORY
                     : (e == hipErrorNoDevice)                                     ? GOOFY_ERR_NO_DEVICE
                                                                                   : GOOFY_ERR_HIP;
    return fail(code, std::string(what) + ": " + hipGetErrorString(e));
}

void clear_error() { t_last_error.clear(); }

// ---- device -------------------------------------------------------------------

int device_info(DeviceInfo* out)
{
    static std::mutex mu;
    static DeviceInfo cache[64];
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        return fail(GOOFY_ERR_NO_DEVICE, "no GPU visible");
    }
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail_hip(e, "hipGetDevice");
    if (dev < 0 || dev >= 64) return fail(GOOFY_ERR_INVALID, "device index out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (cache[dev].device != dev) {
        hipDeviceProp_t prop;
        e = hipGetDeviceProperties(&prop, dev);
        if (e != hipSuccess) return fail_hip(e, "hipGetDeviceProperties");
        cache[dev].num_cus = prop.multiProcessorCount;
        cache[dev].total_mem = prop.totalGlobalMem;
        std::snprintf(cache[dev].arch, sizeof(cache[dev].arch), "%s", prop.gcnArchName);
        cache[dev].device = dev;
    }
    *out = cache[dev];
    return GOOFY_OK;
}

// Measured on MI355X (8192^2, both formats): one block per lane with
// streaming (non-temporal) loads and stores is fastest; 1-4 wave workgroups
// are within noise of each other, 4 blocks per lane loses ~10%.
// One block per lane, streaming, linear tile order, and the largest
// workgroup (8, 4, 2 or 1 waves) that still leaves two workgroups per CU.
// 8-wave workgroups are the fastest shape on large images (on the 8192^2
// image within 1% of the other one-block-per-lane shapes, on 16384^2 images
// with a 65792-byte pitch 20-25% faster than 2-wave ones); a texture of a
// few hundred tiles needs smaller workgroups to reach every CU.
LaunchCfg auto_cfg(uint64_t blocks, int num_cus)
{
    const uint64_t tiles = (blocks + 63u) / 64u;
    const uint64_t want = 2u * (uint64_t)(num_cus > 0 ? num_cus : 1);
    int wg = 8;
    while (wg > 1 && (tiles + (uint64_t)wg - 1) / (uint64_t)wg < want) wg >>= 1;
    return LaunchCfg{1, wg, 0, 1, 0};
}

bool cfg_valid(const LaunchCfg& c)
{
    const bool bpl = c.blocks_per_lane == 1 || c.blocks_per_lane == 2 || c.blocks_per_lane == 4;
    const bool wg = c.wg_waves == 1 || c.wg_waves == 2 || c.wg_waves == 4 || c.wg_waves == 8;
    return bpl && wg && c.waves_per_cu >= 0 && c.waves_per_cu <= 256 && (c.nontemporal == 0 || c.nontemporal == 1) &&
           (c.tile_order == 0 || c.tile_order == 1);
}

// ---- host-memory drop-in: pinned staging, chunked pipeline ---------------------
//
// goofy::compressDXT1/ETC1 take pageable host buffers.  The image is moved in
// chunks of block rows (about 8 MiB of pixels each) through three pinned
// staging slots.  While the GPU copies chunk k in, host threads copy the rows
// of chunk k+1 into the next slot, and the host collects the blocks of chunk
// k-2.  Uploads go on a stream of their own, back to back: the kernel and the
// download of a chunk wait for its upload on the other stream, and a slot is
// refilled only after the host has collected its previous chunk, so the
// upload stream never waits on the device.  (On a single stream each upload
// is ordered behind the previous chunk's kernel and download; on MI355X the
// 8192^2 call took 6.6 ms that way and 5.5 ms with the uploads apart, at the
// link's measured host-to-device ceiling.)  A one-chunk image stays on one
// stream.  Row copies of large chunks are split over a small persistent thread
// pool.  Contexts (stream, events, device and pinned buffers) come from a
// process-wide pool: concurrent callers never share one, and nothing is freed
// at thread exit (the HIP runtime may already be gone then).

constexpr int kSlots = 3;

struct StagingCtx {
    int device = -1;
    hipStream_t stream = nullptr;     // kernels and downloads
    hipStream_t in_stream = nullptr;  // uploads
    hipEvent_t h2d_done[kSlots] = {};   // upload of the slot's chunk finished
    hipEvent_t out_ready[kSlots] = {};  // pinned output slot holds a finished chunk
    void* d_in[kSlots] = {};
    void* d_out[kSlots] = {};
    void* h_in[kSlots] = {};
    void* h_out[kSlots] = {};
    size_t in_cap = 0, out_cap = 0;
};

class StagingPool {
public:
    StagingCtx* acquire(int device)
    {
        std::lock_guard<std::mutex> lock(mu_);
        for (size_t i = 0; i < free_.size(); ++i)
            if (free_[i]->device == device) {
                StagingCtx* c = free_[i];
                free_.erase(free_.begin() + (long)i);
                return c;
            }
        StagingCtx* c = new StagingCtx();
        c->device = device;
        return c;
    }
    void release(StagingCtx* c)
    {
        std::lock_guard<std::mutex> lock(mu_);
        free_.push_back(c);
    }

private:
    std::mutex mu_;
    std::vector<StagingCtx*> free_;
};

static StagingPool& staging_pool()
{
    static StagingPool* pool = new StagingPool();  // never destroyed
    return *pool;
}

static int ensure_ctx(StagingCtx* c, size_t in_bytes, size_t out_bytes)
{
    hipError_t e;
    if (!c->stream) {
        if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess ||
            (e = hipStreamCreateWithFlags(&c->in_stream, hipStreamNonBlocking)) != hipSuccess)
            return fail_hip(e, "hipStreamCreate");
        for (int i = 0; i < kSlots; ++i) {
            if ((e = hipEventCreateWithFlags(&c->h2d_done[i], hipEventDisableTiming)) != hipSuccess ||
                (e = hipEventCreateWithFlags(&c->out_ready[i], hipEventDisableTiming)) != hipSuccess)
                return fail_hip(e, "hipEventCreate");
        }
    }
    if (c->in_cap < in_bytes || c->out_cap < out_bytes) {
        (void)hipStreamSynchronize(c->in_stream);
        (void)hipStreamSynchronize(c->stream);
        const size_t ni = std::max(in_bytes, c->in_cap), no = std::max(out_bytes, c->out_cap);
        for (int i = 0; i < kSlots; ++i) {
            if (c->d_in[i]) (void)hipFree(c->d_in[i]);
            if (c->d_out[i]) (void)hipFree(c->d_out[i]);
            if (c->h_in[i]) (void)hipHostFree(c->h_in[i]);
            if (c->h_out[i]) (void)hipHostFree(c->h_out[i]);
            c->d_in[i] = c->d_out[i] = c->h_in[i] = c->h_out[i] = nullptr;
        }
        c->in_cap = c->out_cap = 0;
        for (int i = 0; i < kSlots; ++i) {
            if ((e = hipMalloc(&c->d_in[i], ni)) != hipSuccess || (e = hipMalloc(&c->d_out[i], no)) != hipSuccess)
                return fail_hip(e, "hipMalloc(staging)");
            if ((e = hipHostMalloc(&c->h_in[i], ni, hipHostMallocDefault)) != hipSuccess ||
                (e = hipHostMalloc(&c->h_out[i], no, hipHostMallocDefault)) != hipSuccess)
                return fail_hip(e, "hipHostMalloc(staging)");
        }
        c->in_cap = ni;
        c->out_cap = no;
    }
    return GOOFY_OK;
}

// Copy `rows` rows of `row_bytes` from src (stride) to dst (packed), split
// over the copy pool when large.
static void copy_rows(uint8_t* dst, const uint8_t* src, size_t stride, size_t row_bytes, unsigned rows)
{
    const size_t total = row_bytes * rows;
    CopyPool& pool = CopyPool::get();
    const int parts = total >= (2u << 20) ? std::min<int>(pool.workers() + 1, (int)std::max<size_t>(1, total >> 20)) : 1;
    if (stride == row_bytes) {
        const size_t step = (total + parts - 1) / parts;
        pool.run(parts, [&](int i) {
            const size_t a = (size_t)i * step, b = std::min(total, a + step);
            if (a < b) std::memcpy(dst + a, src + a, b - a);
        });
        return;
    }
    pool.run(parts, [&](int i) {
        const unsigned r0 = (unsigned)((uint64_t)rows * i / parts), r1 = (unsigned)((uint64_t)rows * (i + 1) / parts);
        for (unsigned r = r0; r < r1; ++r) std::memcpy(dst + r * row_bytes, src + (size_t)r * stride, row_bytes);
    });
}

// fmt: one format into `result`, or kHostDual: DXT1 into `result` and ETC1S
// into `result2` from one pass over the image, or kHostDualAlpha: DXT5 and
// ETC2 RGBA8 the same way.
constexpr int kHostDual = 1 << 16;
constexpr int kHostDualAlpha = (1 << 16) + 1;

static int compress_host(int fmt, unsigned char* result, const unsigned char* input, unsigned width,
                         unsigned height, unsigned stride, unsigned char* result2 = nullptr)
{
    clear_error();
    const bool dual_alpha = fmt == kHostDualAlpha;
    const bool dual = fmt == kHostDual || dual_alpha;
    if (!format_valid(fmt) && !dual) return fail(GOOFY_ERR_INVALID, "unknown format");
    // Same checks, same order as the reference (GoofyTC/goofy_tc.h:1500-1508).
    if (width % 16 != 0) return fail(GOOFY_ERR_WIDTH, "width must be a multiple of 16");
    if (height % 4 != 0) return fail(GOOFY_ERR_HEIGHT, "height must be a multiple of 4");
    if (width == 0 || height == 0) return GOOFY_OK;
    if (!result || !input || (dual && !result2)) return fail(GOOFY_ERR_INVALID, "null buffer");
    if ((size_t)stride < (size_t)width * 4) return fail(GOOFY_ERR_INVALID, "stride smaller than 4*width");
    const uint32_t bw = width / 4, bh = height / 4;
    const uint32_t max_rows = max_stripe_brows(bw);  // per chunk: chunks are launched one by one
    if (max_rows == 0) return fail(GOOFY_ERR_TOO_LARGE, "image too wide (4,194,304 pixels or more per row)");

    DeviceInfo di;
    int rc = device_info(&di);
    if (rc != GOOFY_OK) return rc;
    StagingCtx* ctx = staging_pool().acquire(di.device);
    struct Release {
        StagingCtx* c;
        ~Release() { staging_pool().release(c); }
    } rel{ctx};

    const size_t row_bytes = (size_t)width * 4;
    const unsigned chunk_brows = (unsigned)std::max<size_t>(1, (size_t(8) << 20) / (row_bytes * 4));  // ~8 MiB
    const unsigned cb = std::min(std::min(chunk_brows, bh), max_rows);
    // dual: a chunk's staging output holds its blocks of the first format, then
    // those of the second
    const size_t outs = dual ? 2 : 1;
    const size_t bb = dual_alpha ? 16 : (dual ? 8 : block_bytes(fmt));
    if ((rc = ensure_ctx(ctx, row_bytes * 4 * cb, outs * bw * cb * bb)) != GOOFY_OK) return rc;
    const unsigned nchunks = (bh + cb - 1) / cb;
    const LaunchCfg cfg = auto_cfg((uint64_t)bw * cb, di.num_cus);
    hipError_t e = hipSuccess;

    auto finish_chunk = [&](unsigned k) -> hipError_t {
        const int s = (int)(k % kSlots);
        const unsigned rows = std::min(cb, bh - k * cb);
        hipError_t r = hipEventSynchronize(ctx->out_ready[s]);
        if (r != hipSuccess) return r;
        const size_t n = (size_t)rows * bw * bb;
        std::memcpy(result + (size_t)k * cb * bw * bb, ctx->h_out[s], n);
        if (dual) std::memcpy(result2 + (size_t)k * cb * bw * bb, static_cast<uint8_t*>(ctx->h_out[s]) + n, n);
        return hipSuccess;
    };

    // Chunk k uses slot k % kSlots.  Chunk k-kSlots, the slot's previous user,
    // was collected by finish_chunk in an earlier iteration, so its upload,
    // kernel and download are complete and the slot is free on both sides.
    // An image of one chunk (up to about 8 MiB) is not copied to the device:
    // the kernels read the pinned staging copy and write the pinned output
    // over PCIe themselves (zero-copy), which saves the two DMA transfers'
    // set-up and synchronisation -- at 512x512, 66 against 82 us.
    const bool two_streams = nchunks > 1;
    const bool zero_copy = nchunks == 1;
    if (zero_copy) {
        // The host stages the image in up to 8 parts of at least 512 KiB, and
        // each part's kernel is launched as soon as its rows are staged: the
        // GPU reads part p over PCIe while the host copies part p+1.  One
        // stream, one synchronisation (finish_chunk).  At 512x512, 2 parts took
        // 0.052-0.054 ms, 4 parts of 256 KiB 0.054-0.055 and 8 parts 0.064-0.066
        // (each launch costs the host a few microseconds); at 1024x1024, 8
        // parts took 0.145 ms against 0.151 for 4.  The kernels read PCIe at
        // about 35 GB/s against the copy engines' 41-55, so larger images keep
        // the DMA pipeline (at 2896x2896 zero-copy took 1.01 ms against 0.81).
        uint8_t* const hin = static_cast<uint8_t*>(ctx->h_in[0]);
        uint8_t* const hout = static_cast<uint8_t*>(ctx->h_out[0]);
        const size_t n = (size_t)bw * bh * bb;  // bytes of one format's blocks
        const unsigned parts = (unsigned)std::min<size_t>(8, std::max<size_t>(1, (row_bytes * 4 * bh) >> 19));
        const unsigned prows = (bh + parts - 1) / parts;
        for (unsigned r0 = 0; r0 < bh && e == hipSuccess; r0 += prows) {
            const unsigned rows = std::min(prows, bh - r0);
            uint8_t* const in = hin + (size_t)r0 * 4 * row_bytes;
            uint8_t* const out = hout + (size_t)r0 * bw * bb;
            copy_rows(in, input + (size_t)r0 * 4 * stride, stride, row_bytes, rows * 4);
            e = dual_alpha ? launch_dual_alpha((uint32_t)cfg.wg_waves, in, row_bytes, out, out + n, bw, bw * rows,
                                               ctx->stream)
                : dual     ? launch_dual((uint32_t)cfg.wg_waves, in, row_bytes, out, out + n, bw, bw * rows, ctx->stream)
                           : launch_plain(fmt, cfg, in, row_bytes, out, bw, bw * rows, ctx->stream);
        }
        if (e == hipSuccess) e = hipEventRecord(ctx->out_ready[0], ctx->stream);
        if (e == hipSuccess) e = finish_chunk(0);
        if (e != hipSuccess) {
            (void)hipStreamSynchronize(ctx->stream);
            return fail_hip(e, "drop-in zero-copy launch");
        }
        return GOOFY_OK;
    }
    hipStream_t up = two_streams ? ctx->in_stream : ctx->stream;
    unsigned collected = 0;  // chunks [0, collected) are in `result`
    for (unsigned k = 0; k < nchunks && e == hipSuccess; ++k) {
        const int s = (int)(k % kSlots);
        const unsigned rows = std::min(cb, bh - k * cb);  // block rows in this chunk
        copy_rows(static_cast<uint8_t*>(ctx->h_in[s]), input + (size_t)k * cb * 4 * stride, stride, row_bytes,
                  rows * 4);
        void* const kin = ctx->d_in[s];
        void* const kout = ctx->d_out[s];
        if ((e = hipMemcpyAsync(ctx->d_in[s], ctx->h_in[s], row_bytes * 4 * rows, hipMemcpyHostToDevice, up)) !=
            hipSuccess)
            break;
        if (two_streams) {
            if ((e = hipEventRecord(ctx->h2d_done[s], up)) != hipSuccess) break;
            if ((e = hipStreamWaitEvent(ctx->stream, ctx->h2d_done[s], 0)) != hipSuccess) break;
        }
        const size_t n = (size_t)bw * rows * bb;
        uint8_t* second = static_cast<uint8_t*>(kout) + n;
        e = dual_alpha ? launch_dual_alpha((uint32_t)cfg.wg_waves, kin, row_bytes, kout, second, bw, bw * rows,
                                           ctx->stream)
            : dual     ? launch_dual((uint32_t)cfg.wg_waves, kin, row_bytes, kout, second, bw, bw * rows, ctx->stream)
                       : launch_plain(fmt, cfg, kin, row_bytes, kout, bw, bw * rows, ctx->stream);
        if (e != hipSuccess) break;
        if ((e = hipMemcpyAsync(ctx->h_out[s], ctx->d_out[s], outs * n, hipMemcpyDeviceToHost, ctx->stream)) !=
            hipSuccess)
            break;
        if ((e = hipEventRecord(ctx->out_ready[s], ctx->stream)) != hipSuccess) break;
        // collect what is kSlots - 1 chunks behind, freeing the slot chunk k+1 uses
        while (e == hipSuccess && collected + (kSlots - 1) <= k) e = finish_chunk(collected++);
    }
    while (e == hipSuccess && collected < nchunks) e = finish_chunk(collected++);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(ctx->in_stream);
        (void)hipStreamSynchronize(ctx->stream);
        return fail_hip(e, "drop-in staging pipeline");
    }
    return GOOFY_OK;
}

static int compress_device(int fmt, void* d_result, const void* d_input, unsigned width, unsigned height,
                           size_t pitch, void* stream, const goofy_launch_config* ucfg, void* ev_start = nullptr,
                           void* ev_stop = nullptr)
{
    clear_error();
    if (!format_valid(fmt)) return fail(GOOFY_ERR_INVALID, "unknown format");
    if (width % 4 != 0) return fail(GOOFY_ERR_WIDTH, "width must be a multiple of 4");
    if (height % 4 != 0) return fail(GOOFY_ERR_HEIGHT, "height must be a multiple of 4");
    if (width == 0 || height == 0) return GOOFY_OK;
    if (!d_result || !d_input) return fail(GOOFY_ERR_INVALID, "null buffer");
    if (pitch < (size_t)width * 4 || pitch % 16 != 0)
        return fail(GOOFY_ERR_INVALID, "pitch must be >= 4*width and a multiple of 16");
    const size_t bb = block_bytes(fmt);
    if ((uintptr_t)d_input % 16 != 0 || (uintptr_t)d_result % bb != 0)
        return fail(GOOFY_ERR_INVALID, "input must be 16-byte and result block-size (8 or 16 byte) aligned");
    const uint32_t bw = width / 4, bh = height / 4;
    const uint32_t max_rows = max_stripe_brows(bw);
    if (max_rows == 0) return fail(GOOFY_ERR_TOO_LARGE, "image too wide (4,194,304 pixels or more per row)");
    LaunchCfg cfg{};
    int num_cus = 0;
    if (ucfg) {
        cfg = to_cfg(*ucfg);
        if (!cfg_valid(cfg)) return fail(GOOFY_ERR_INVALID, "invalid launch config");
    } else {
        DeviceInfo di;
        int rc = device_info(&di);
        if (rc != GOOFY_OK) return rc;
        num_cus = di.num_cus;
    }
    // One launch, or horizontal stripes for images beyond one launch's index
    // range; the start event goes with the first launch, the stop event with
    // the last, so the interval covers them all.
    for (uint32_t y0 = 0; y0 < bh;) {
        const uint32_t rows = std::min(max_rows, bh - y0);
        const uint32_t nb = rows * bw;  // <= 0xFFFFFFFF by max_stripe_brows
        const LaunchCfg c = ucfg ? cfg : auto_cfg(nb, num_cus);
        hipError_t e = launch_plain(fmt, c, static_cast<const uint8_t*>(d_input) + (size_t)y0 * 4 * pitch, pitch,
                                    static_cast<uint8_t*>(d_result) + (size_t)y0 * bw * bb, bw, nb,
                                    (hipStream_t)stream, y0 == 0 ? (hipEvent_t)ev_start : nullptr,
                                    y0 + rows == bh ? (hipEvent_t)ev_stop : nullptr);
        if (e != hipSuccess) return fail_hip(e, "encode kernel launch");
        y0 += rows;
    }
    return GOOFY_OK;
}

}  // namespace goofy_impl

using namespace goofy_impl;

extern "C" {

const char* goofy_hip_last_error(void) { return t_last_error.c_str(); }

const char* goofy_hip_status_string(int s)
{
    switch (s) {
    case GOOFY_OK: return "ok";
    case GOOFY_ERR_WIDTH: return "bad width";
    case GOOFY_ERR_HEIGHT: return "bad height";
    case GOOFY_ERR_HIP: return "HIP runtime error";
    case GOOFY_ERR_INVALID: return "invalid argument";
    case GOOFY_ERR_OUT_OF_MEMORY: return "out of memory";
    case GOOFY_ERR_NO_DEVICE: return "no GPU";
    case GOOFY_ERR_TOO_LARGE: return "image too large";
    case GOOFY_ERR_BUSY: return "busy";
    case GOOFY_ERR_CANCELLED: return "cancelled";
    case GOOFY_ERR_DEADLINE: return "deadline passed";
    default: return "unknown status";
    }
}

const char* goofy_hip_version(void) { return "goofy_hip 0.1 (gfx950)"; }

int goofy_hip_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

int goofy_hip_compress(int format, unsigned char* result, const unsigned char* input, unsigned int width,
                       unsigned int height, unsigned int stride)
{
    return compress_host(format, result, input, width, height, stride);
}

int goofy_hip_compress_dxt1(unsigned char* result, const unsigned char* input, unsigned int width, unsigned int height,
                            unsigned int stride)
{
    return compress_host(GOOFY_FORMAT_DXT1, result, input, width, height, stride);
}

int goofy_hip_compress_etc1(unsigned char* result, const unsigned char* input, unsigned int width, unsigned int height,
                            unsigned int stride)
{
    return compress_host(GOOFY_FORMAT_ETC1, result, input, width, height, stride);
}

int goofy_hip_compress_dxt5(unsigned char* result, const unsigned char* input, unsigned int width, unsigned int height,
                            unsigned int stride)
{
    return compress_host(GOOFY_FORMAT_DXT5, result, input, width, height, stride);
}

int goofy_hip_compress_etc2(unsigned char* result, const unsigned char* input, unsigned int width, unsigned int height,
                            unsigned int stride)
{
    return compress_host(GOOFY_FORMAT_ETC2, result, input, width, height, stride);
}

unsigned int goofy_format_block_bytes(int format) { return format_valid(format) ? (unsigned)block_bytes(format) : 0u; }

int goofy_hip_compress_device(int format, void* d_result, const void* d_input, unsigned int width,
                              unsigned int height, size_t pitch, void* stream)
{
    return compress_device(format, d_result, d_input, width, height, pitch, stream, nullptr);
}

int goofy_hip_compress_device_cfg(int format, void* d_result, const void* d_input, unsigned int width,
                                  unsigned int height, size_t pitch, void* stream, const goofy_launch_config* cfg)
{
    return compress_device(format, d_result, d_input, width, height, pitch, stream, cfg);
}

int goofy_hip_compress_device_timed(int format, void* d_result, const void* d_input, unsigned int width,
                                    unsigned int height, size_t pitch, void* stream, const goofy_launch_config* cfg,
                                    void* start_event, void* stop_event)
{
    return compress_device(format, d_result, d_input, width, height, pitch, stream, cfg, start_event, stop_event);
}

int goofy_hip_compress_dual(unsigned char* result_dxt1, unsigned char* result_etc1, const unsigned char* input,
                            unsigned int width, unsigned int height, unsigned int stride)
{
    return compress_host(kHostDual, result_dxt1, input, width, height, stride, result_etc1);
}

static int compress_device_pair(bool alpha, void* d_a, void* d_b, const void* d_input, unsigned int width,
                                unsigned int height, size_t pitch, void* stream, void* start_event, void* stop_event)
{
    clear_error();
    if (width % 4 != 0) return fail(GOOFY_ERR_WIDTH, "width must be a multiple of 4");
    if (height % 4 != 0) return fail(GOOFY_ERR_HEIGHT, "height must be a multiple of 4");
    if (width == 0 || height == 0) return GOOFY_OK;
    if (!d_a || !d_b || !d_input) return fail(GOOFY_ERR_INVALID, "null buffer");
    if (pitch < (size_t)width * 4 || pitch % 16 != 0)
        return fail(GOOFY_ERR_INVALID, "pitch must be >= 4*width and a multiple of 16");
    const size_t bb = alpha ? 16 : 8;
    if ((uintptr_t)d_input % 16 != 0 || (uintptr_t)d_a % bb != 0 || (uintptr_t)d_b % bb != 0)
        return fail(GOOFY_ERR_INVALID, "input must be 16-byte and results block-size (8 or 16 byte) aligned");
    const uint32_t bw = width / 4, bh = height / 4;
    const uint32_t max_rows = max_stripe_brows(bw);
    if (max_rows == 0) return fail(GOOFY_ERR_TOO_LARGE, "image too wide (4,194,304 pixels or more per row)");
    DeviceInfo di;
    int rc = device_info(&di);
    if (rc != GOOFY_OK) return rc;
    for (uint32_t y0 = 0; y0 < bh;) {  // stripes for images beyond one launch's index range
        const uint32_t rows = std::min(max_rows, bh - y0);
        const uint32_t nb = rows * bw;
        const size_t out_off = (size_t)y0 * bw * bb;
        const uint32_t wg = (uint32_t)auto_cfg(nb, di.num_cus).wg_waves;
        const uint8_t* in = static_cast<const uint8_t*>(d_input) + (size_t)y0 * 4 * pitch;
        uint8_t* a = static_cast<uint8_t*>(d_a) + out_off;
        uint8_t* b = static_cast<uint8_t*>(d_b) + out_off;
        hipEvent_t e0 = y0 == 0 ? (hipEvent_t)start_event : nullptr;
        hipEvent_t e1 = y0 + rows == bh ? (hipEvent_t)stop_event : nullptr;
        hipError_t e = alpha ? launch_dual_alpha(wg, in, pitch, a, b, bw, nb, (hipStream_t)stream, e0, e1)
                             : launch_dual(wg, in, pitch, a, b, bw, nb, (hipStream_t)stream, e0, e1);
        if (e != hipSuccess) return fail_hip(e, "dual encode kernel launch");
        y0 += rows;
    }
    return GOOFY_OK;
}

int goofy_hip_compress_device_dual(void* d_dxt1, void* d_etc1, const void* d_input, unsigned int width,
                                   unsigned int height, size_t pitch, void* stream, void* start_event,
                                   void* stop_event)
{
    return compress_device_pair(false, d_dxt1, d_etc1, d_input, width, height, pitch, stream, start_event,
                                stop_event);
}

int goofy_hip_compress_device_dual_alpha(void* d_dxt5, void* d_etc2, const void* d_input, unsigned int width,
                                         unsigned int height, size_t pitch, void* stream, void* start_event,
                                         void* stop_event)
{
    return compress_device_pair(true, d_dxt5, d_etc2, d_input, width, height, pitch, stream, start_event,
                                stop_event);
}

int goofy_hip_compress_dual_alpha(unsigned char* result_dxt5, unsigned char* result_etc2, const unsigned char* input,
                                  unsigned int width, unsigned int height, unsigned int stride)
{
    return compress_host(kHostDualAlpha, result_dxt5, input, width, height, stride, result_etc2);
}

int goofy_hip_tune_device(int format, void* d_result, const void* d_input, unsigned int width, unsigned int height,
                          size_t pitch, void* stream, int reps, goofy_launch_config* best, float* best_ms)
{
    // validates the arguments and warms the device with the default shape
    int rc = compress_device(format, d_result, d_input, width, height, pitch, stream, nullptr);
    if (rc != GOOFY_OK) return rc;
    if (width == 0 || height == 0) return fail(GOOFY_ERR_INVALID, "empty image");
    if (reps < 1) reps = 1;
    hipEvent_t ev[2] = {nullptr, nullptr};
    for (hipEvent_t& x : ev) {
        hipError_t e = hipEventCreate(&x);
        if (e != hipSuccess) {
            for (hipEvent_t y : ev)
                if (y) (void)hipEventDestroy(y);
            return fail_hip(e, "hipEventCreate");
        }
    }
    LaunchCfg top{1, 8, 0, 1, 0};
    float top_ms = 1e30f;
    const int bpls[] = {1, 2, 4}, wgs[] = {1, 2, 4, 8}, orders[] = {0, 1};
    for (int bpl : bpls)
        for (int wg : wgs)
            for (int order : orders) {
                const goofy_launch_config c = to_public(LaunchCfg{bpl, wg, 0, 1, order});
                std::vector<float> ts;
                for (int r = 0; r <= reps && rc == GOOFY_OK; ++r) {  // r == 0: warm-up
                    rc = compress_device(format, d_result, d_input, width, height, pitch, stream, &c, ev[0], ev[1]);
                    hipError_t e = rc == GOOFY_OK ? hipEventSynchronize(ev[1]) : hipSuccess;
                    if (e != hipSuccess) rc = fail_hip(e, "hipEventSynchronize");
                    float ms = 0;
                    if (rc == GOOFY_OK && r > 0 && hipEventElapsedTime(&ms, ev[0], ev[1]) == hipSuccess)
                        ts.push_back(ms);
                }
                if (rc != GOOFY_OK) break;
                if (ts.empty()) continue;
                std::sort(ts.begin(), ts.end());
                if (ts[ts.size() / 2] < top_ms) {
                    top_ms = ts[ts.size() / 2];
                    top = to_cfg(c);
                }
            }
    for (hipEvent_t y : ev) (void)hipEventDestroy(y);
    if (rc != GOOFY_OK) return rc;
    if (best) *best = to_public(top);
    if (best_ms) *best_ms = top_ms;
    return GOOFY_OK;
}

int goofy_synth_fill_host(int kind, uint64_t seed, uint32_t texture_id, uint32_t noise, unsigned char* dst,
                          unsigned int width, unsigned int height, size_t pitch, int threads)
{
    clear_error();
    if (kind < 0 || kind > 2) return fail(GOOFY_ERR_INVALID, "unknown synth kind");
    if (width == 0 || height == 0) return GOOFY_OK;
    if (!dst || pitch < (size_t)width * 4) return fail(GOOFY_ERR_INVALID, "bad destination");
    const goofy_synth::Params P = goofy_synth::make_params((uint32_t)kind, seed, texture_id, noise);
    if (threads <= 0) threads = (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    threads = (int)std::min<unsigned>((unsigned)threads, height);
    auto rows = [&](unsigned y0, unsigned y1) {
        for (unsigned y = y0; y < y1; ++y) {
            uint32_t* row = reinterpret_cast<uint32_t*>(dst + (size_t)y * pitch);
            for (unsigned x = 0; x < width; ++x) {
                const uint32_t v = goofy_synth::pixel(P, x, y);
                std::memcpy(row + x, &v, 4);
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t)
        pool.emplace_back(rows, (unsigned)((uint64_t)height * t / threads), (unsigned)((uint64_t)height * (t + 1) / threads));
    rows(0, (unsigned)((uint64_t)height / threads));
    for (auto& th : pool) th.join();
    return GOOFY_OK;
}

int goofy_synth_fill_device(int kind, uint64_t seed, uint32_t texture_id, uint32_t noise, void* d_dst,
                            unsigned int width, unsigned int height, size_t pitch, void* stream)
{
    clear_error();
    if (kind < 0 || kind > 2) return fail(GOOFY_ERR_INVALID, "unknown synth kind");
    if (width == 0 || height == 0) return GOOFY_OK;
    if (!d_dst || pitch < (size_t)width * 4 || pitch % 4 != 0 || (uintptr_t)d_dst % 4 != 0)
        return fail(GOOFY_ERR_INVALID, "bad destination");
    if (height > 65535u) return fail(GOOFY_ERR_TOO_LARGE, "height above 65535");
    const goofy_synth::Params P = goofy_synth::make_params((uint32_t)kind, seed, texture_id, noise);
    hipError_t e = launch_synth(&P, d_dst, width, height, pitch, (hipStream_t)stream);
    if (e != hipSuccess) return fail_hip(e, "synth kernel launch");
    return GOOFY_OK;
}

}  // extern "C"

// The drop-in C++ API (include/goofy_tc.h).
namespace goofy {
__attribute__((visibility("default"))) int compressDXT1(unsigned char* result, const unsigned char* input, unsigned int width, unsigned int height,
                 unsigned int stride)
{
    return goofy_impl::compress_host(GOOFY_FORMAT_DXT1, result, input, width, height, stride);
}
__attribute__((visibility("default"))) int compressETC1(unsigned char* result, const unsigned char* input,
                                                        unsigned int width, unsigned int height, unsigned int stride)
{
    return goofy_impl::compress_host(GOOFY_FORMAT_ETC1, result, input, width, height, stride);
}
__attribute__((visibility("default"))) int compressDXT5(unsigned char* result, const unsigned char* input,
                                                        unsigned int width, unsigned int height, unsigned int stride)
{
    return goofy_impl::compress_host(GOOFY_FORMAT_DXT5, result, input, width, height, stride);
}
__attribute__((visibility("default"))) int compressETC2(unsigned char* result, const unsigned char* input,
                                                        unsigned int width, unsigned int height, unsigned int stride)
{
    return goofy_impl::compress_host(GOOFY_FORMAT_ETC2, result, input, width, height, stride);
}
}  // namespace goofy
