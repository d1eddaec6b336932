// codec_api.cpp -- C-ABI of the kernels either side of the encoder: the
// decoders, the round-trip and image-against-image quality sums with the
// reference harness's MSE / PSNR definitions (getMsePsnr, Src/main.cpp:403-469),
// and the harness's RGB565 quality floor (Src/main.cpp:501-555).  The kernels
// are in codec_extra.hip.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "goofy_hip.h"
#include "internal.h"

using namespace goofy_impl;

extern "C" {

// striped = the caller launches stripes of max_stripe_brows block rows, so
// only a single block row has to fit one launch's index range.
static int check_image_device(int fmt, const void* d_img, const void* d_blocks, unsigned width, unsigned height,
                              size_t pitch, bool striped)
{
    if (!format_valid(fmt)) return fail(GOOFY_ERR_INVALID, "unknown format");
    if (width % 4 != 0) return fail(GOOFY_ERR_WIDTH, "width must be a multiple of 4");
    if (height % 4 != 0) return fail(GOOFY_ERR_HEIGHT, "height must be a multiple of 4");
    if (width == 0 || height == 0) return GOOFY_OK;
    if (!d_img || !d_blocks) return fail(GOOFY_ERR_INVALID, "null buffer");
    if (pitch < (size_t)width * 4 || pitch % 16 != 0)
        return fail(GOOFY_ERR_INVALID, "pitch must be >= 4*width and a multiple of 16");
    if ((uintptr_t)d_img % 16 != 0 || (uintptr_t)d_blocks % block_bytes(fmt) != 0)
        return fail(GOOFY_ERR_INVALID, "image must be 16-byte and blocks block-size (8 or 16 byte) aligned");
    const uint64_t bw = width / 4, nb = bw * (height / 4);
    if (striped ? max_stripe_brows(bw) == 0 : (!magic_ok(bw, nb) || nb > 0xFFFFFFFFull))
        return fail(GOOFY_ERR_TOO_LARGE, "image too large");
    return GOOFY_OK;
}

int goofy_hip_decode_device(int format, void* d_rgba, const void* d_blocks, unsigned int width, unsigned int height,
                            size_t pitch, void* stream)
{
    clear_error();
    int rc = check_image_device(format, d_rgba, d_blocks, width, height, pitch, true);
    if (rc != GOOFY_OK || width == 0 || height == 0) return rc;
    const uint32_t bw = width / 4, bh = height / 4, max_rows = max_stripe_brows(bw);
    for (uint32_t y0 = 0; y0 < bh;) {  // stripes for images beyond one launch's index range
        const uint32_t rows = std::min(max_rows, bh - y0);
        hipError_t e = launch_decode(format, static_cast<const uint8_t*>(d_blocks) + (size_t)y0 * bw * block_bytes(format),
                                     static_cast<uint8_t*>(d_rgba) + (size_t)y0 * 4 * pitch, width, rows * 4, pitch,
                                     (hipStream_t)stream);
        if (e != hipSuccess) return fail_hip(e, "decode kernel launch");
        y0 += rows;
    }
    return GOOFY_OK;
}

static double psnr_of(double mse, double peak)
{
    if (mse < 0.0000001) return DBL_MAX;
    return 10.0 * std::log10((peak * peak) / mse);
}

}  // extern "C"

void goofy_impl::quality_figures(const QualitySums& h, uint64_t pixels, goofy_quality* out)
{
    goofy_quality q{};
    q.pixels = pixels;
    for (int i = 0; i < 4; ++i) q.sse_byte[i] = h.e[i];
    q.sse_luma = (double)h.y100 / 10000.0;
    const double n = (double)q.pixels;
    // labels as in the reference harness: "r" = byte 2, "b" = byte 0
    q.mse_r = (double)q.sse_byte[2] / n;
    q.mse_g = (double)q.sse_byte[1] / n;
    q.mse_b = (double)q.sse_byte[0] / n;
    q.mse_a = (double)q.sse_byte[3] / n;
    q.mse_rgb = (double)(q.sse_byte[0] + q.sse_byte[1] + q.sse_byte[2]) / n;
    q.mse_y = q.sse_luma / n;
    q.mse_max = std::max(std::max(q.mse_r, q.mse_g), std::max(q.mse_b, q.mse_a));
    q.psnr_r = psnr_of(q.mse_r, 255.0);
    q.psnr_g = psnr_of(q.mse_g, 255.0);
    q.psnr_b = psnr_of(q.mse_b, 255.0);
    q.psnr_a = psnr_of(q.mse_a, 255.0);
    q.psnr_min = std::min(std::min(q.psnr_r, q.psnr_g), std::min(q.psnr_b, q.psnr_a));
    q.psnr_rgb = psnr_of(q.mse_rgb, 768.0);
    q.psnr_y = psnr_of(q.mse_y, 255.0);
    *out = q;
}

extern "C" {

// Device scratch of the quality sums, kept per device for reuse.  A call
// returns its buffer only after synchronising its stream, so a buffer is never
// shared by two calls in flight; no call frees one, because hipFree would wait
// for every stream on the device (a quality check during serving would stall
// the server's streams).
// A quality call's buffers: device scratch (quality_scratch_bytes) and one
// QualitySums of pinned host memory that the final-sum kernel writes.
struct QualitySlot {
    int device;
    void* d;
    void* h;
};

class QualityScratchPool {
public:
    static constexpr size_t kBytes = kQualityNoAllocBytes;

    // A free slot of `device`, or a new one (device scratch and pinned host
    // memory allocated).  d == nullptr on failure.
    QualitySlot acquire(int device)
    {
        {
            std::lock_guard<std::mutex> lock(mu_);
            for (size_t i = 0; i < free_.size(); ++i)
                if (free_[i].device == device) {
                    const QualitySlot s = free_[i];
                    free_.erase(free_.begin() + (long)i);
                    return s;
                }
        }
        QualitySlot s{device, nullptr, nullptr};
        if (hipMalloc(&s.d, kBytes) != hipSuccess) return {device, nullptr, nullptr};
        if (hipHostMalloc(&s.h, sizeof(QualitySums), hipHostMallocDefault) != hipSuccess) {
            (void)hipFree(s.d);
            return {device, nullptr, nullptr};
        }
        return s;
    }
    void release(const QualitySlot& s)
    {
        std::lock_guard<std::mutex> lock(mu_);
        free_.push_back(s);
    }

private:
    std::mutex mu_;
    std::vector<QualitySlot> free_;
};

static QualityScratchPool& quality_scratch_pool()
{
    static QualityScratchPool* pool = new QualityScratchPool();  // never destroyed
    return *pool;
}

// Runs `launch(scratch, host)` -- quality launches on `stream` whose last one
// leaves the sums at offset 0 of scratch and at host (pinned) -- waits for the
// stream and fills `out` with getMsePsnr's figures (Src/main.cpp:403-469).
extern "C++" template <typename Launch>
static int run_quality(uint64_t pixels, uint64_t units, hipStream_t stream, goofy_quality* out, Launch&& launch)
{
    if (quality_scratch_bytes(units) > QualityScratchPool::kBytes) return fail(GOOFY_ERR_TOO_LARGE, "quality grid");
    DeviceInfo di;
    int rc = device_info(&di);
    if (rc != GOOFY_OK) return rc;
    const QualitySlot slot = quality_scratch_pool().acquire(di.device);
    if (!slot.d) return fail(GOOFY_ERR_OUT_OF_MEMORY, "quality scratch");
    struct Release {  // on every path, the stream has drained before the slot is reused
        QualitySlot slot;
        hipStream_t s;
        ~Release()
        {
            (void)hipStreamSynchronize(s);
            quality_scratch_pool().release(slot);
        }
    } guard{slot, stream};
    hipError_t e = launch(slot.d, slot.h);
    if (e != hipSuccess) return fail_hip(e, "quality kernel launch");
    e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return fail_hip(e, "quality readback");
    QualitySums h;
    std::memcpy(&h, slot.h, sizeof h);
    quality_figures(h, pixels, out);
    return GOOFY_OK;
}

int goofy_hip_quality_device(int format, const void* d_blocks, const void* d_source, unsigned int width,
                             unsigned int height, size_t pitch, void* stream, goofy_quality* out)
{
    clear_error();
    if (!out) return fail(GOOFY_ERR_INVALID, "null output");
    int rc = check_image_device(format, d_source, d_blocks, width, height, pitch, true);
    if (rc != GOOFY_OK) return rc;
    if (width == 0 || height == 0) return fail(GOOFY_ERR_INVALID, "empty image");
    const uint32_t bw = width / 4, bh = height / 4, max_rows = quality_stripe_rows(bw);
    const uint64_t stripe_blocks = (uint64_t)bw * std::min(max_rows, bh);
    return run_quality((uint64_t)width * height, stripe_blocks, (hipStream_t)stream, out, [&](void* scratch, void* host) {
        hipError_t e = hipSuccess;
        for (uint32_t y0 = 0; y0 < bh && e == hipSuccess;) {  // stripes beyond one launch's range
            const uint32_t rows = std::min(max_rows, bh - y0);
            e = launch_sqerr(format, static_cast<const uint8_t*>(d_blocks) + (size_t)y0 * bw * block_bytes(format),
                             static_cast<const uint8_t*>(d_source) + (size_t)y0 * 4 * pitch, pitch, width, rows * 4,
                             scratch, y0 > 0, host, (hipStream_t)stream);
            y0 += rows;
        }
        return e;
    });
}

// As goofy_hip_quality_device with the library's fixed device array (one per
// GPU, kQualityNoAllocBytes) as the work area, and the sums read back with a
// plain copy into the caller's stack: no device or pinned allocation.  Calls
// that overlap on one GPU share the array and race.
int goofy_hip_quality_device_noalloc(int format, const void* d_blocks, const void* d_source, unsigned int width,
                                     unsigned int height, size_t pitch, void* stream, goofy_quality* out)
{
    clear_error();
    if (!out) return fail(GOOFY_ERR_INVALID, "null output");
    int rc = check_image_device(format, d_source, d_blocks, width, height, pitch, true);
    if (rc != GOOFY_OK) return rc;
    if (width == 0 || height == 0) return fail(GOOFY_ERR_INVALID, "empty image");
    const uint32_t bw = width / 4, bh = height / 4, max_rows = quality_stripe_rows(bw);
    const uint64_t stripe_blocks = (uint64_t)bw * std::min(max_rows, bh);
    if (quality_scratch_bytes(stripe_blocks) > kQualityNoAllocBytes)
        return fail(GOOFY_ERR_TOO_LARGE, "quality grid larger than the fixed work area");
    DeviceInfo di;
    if ((rc = device_info(&di)) != GOOFY_OK) return rc;
    void* work = nullptr;
    hipError_t e = noalloc_scratch_address(di.device, &work);
    if (e != hipSuccess) return fail_hip(e, "hipGetSymbolAddress(quality work area)");
    const hipStream_t s = (hipStream_t)stream;
    for (uint32_t y0 = 0; y0 < bh && e == hipSuccess;) {  // stripes beyond one launch's range
        const uint32_t rows = std::min(max_rows, bh - y0);
        e = launch_sqerr(format, static_cast<const uint8_t*>(d_blocks) + (size_t)y0 * bw * block_bytes(format),
                         static_cast<const uint8_t*>(d_source) + (size_t)y0 * 4 * pitch, pitch, width, rows * 4, work,
                         y0 > 0, nullptr, s);
        y0 += rows;
    }
    if (e != hipSuccess) return fail_hip(e, "quality kernel launch");
    QualitySums h;
    if ((e = hipMemcpyAsync(&h, work, sizeof h, hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (e = hipStreamSynchronize(s)) != hipSuccess)
        return fail_hip(e, "quality readback");
    quality_figures(h, (uint64_t)width * height, out);
    return GOOFY_OK;
}

// Checks shared by the two image-to-image entry points: width % 4 (one lane
// per 4 pixels), 16-byte aligned rows, and sizes the kernels' index
// arithmetic covers.
static int check_rgba_pair(const void* d_a, size_t pitch_a, const void* d_b, size_t pitch_b, unsigned width,
                           unsigned height, unsigned height_multiple)
{
    if (width % 4 != 0) return fail(GOOFY_ERR_WIDTH, "width must be a multiple of 4");
    if (height % height_multiple != 0)
        return fail(GOOFY_ERR_HEIGHT, height_multiple == 2 ? "height must be even" : "bad height");
    if (width == 0 || height == 0) return GOOFY_OK;
    if (!d_a || !d_b) return fail(GOOFY_ERR_INVALID, "null buffer");
    if (pitch_a < (size_t)width * 4 || pitch_a % 16 != 0 || pitch_b < (size_t)width * 4 || pitch_b % 16 != 0)
        return fail(GOOFY_ERR_INVALID, "pitches must be >= 4*width and multiples of 16");
    if ((uintptr_t)d_a % 16 != 0 || (uintptr_t)d_b % 16 != 0)
        return fail(GOOFY_ERR_INVALID, "images must be 16-byte aligned");
    if (max_stripe_brows(width / 4) == 0)  // larger images run as stripes of rows
        return fail(GOOFY_ERR_TOO_LARGE, "image too wide (4,194,304 pixels or more per row)");
    return GOOFY_OK;
}

int goofy_hip_quality_images_device(const void* d_a, size_t pitch_a, const void* d_b, size_t pitch_b,
                                    unsigned int width, unsigned int height, void* stream, goofy_quality* out)
{
    clear_error();
    if (!out) return fail(GOOFY_ERR_INVALID, "null output");
    int rc = check_rgba_pair(d_a, pitch_a, d_b, pitch_b, width, height, 1);
    if (rc != GOOFY_OK) return rc;
    if (width == 0 || height == 0) return fail(GOOFY_ERR_INVALID, "empty image");
    const uint32_t qw = width / 4, max_rows = quality_stripe_rows(qw);  // rows of qw 4-pixel units
    const uint64_t stripe_units = (uint64_t)qw * std::min(max_rows, height);
    return run_quality((uint64_t)width * height, stripe_units, (hipStream_t)stream, out, [&](void* scratch, void* host) {
        hipError_t e = hipSuccess;
        for (uint32_t y0 = 0; y0 < height && e == hipSuccess;) {
            const uint32_t rows = std::min(max_rows, height - y0);
            e = launch_sqerr_images(static_cast<const uint8_t*>(d_a) + (size_t)y0 * pitch_a, pitch_a,
                                    static_cast<const uint8_t*>(d_b) + (size_t)y0 * pitch_b, pitch_b, width, rows,
                                    scratch, y0 > 0, host, (hipStream_t)stream);
            y0 += rows;
        }
        return e;
    });
}

int goofy_hip_rgb565_baseline_device(void* d_dst, size_t dst_pitch, const void* d_src, size_t src_pitch,
                                     unsigned int width, unsigned int height, void* stream)
{
    return goofy_hip_rgb565_baseline_device_timed(d_dst, dst_pitch, d_src, src_pitch, width, height, stream, nullptr,
                                                  nullptr);
}

int goofy_hip_rgb565_baseline_device_timed(void* d_dst, size_t dst_pitch, const void* d_src, size_t src_pitch,
                                           unsigned int width, unsigned int height, void* stream, void* start_event,
                                           void* stop_event)
{
    clear_error();
    int rc = check_rgba_pair(d_dst, dst_pitch, d_src, src_pitch, width, height, 2);
    if (rc != GOOFY_OK || width == 0 || height == 0) return rc;
    // the kernel's rows are pairs of pixel rows; stripes of at most
    // max_stripe_brows such pairs keep every launch within its index range
    const uint32_t max_px_rows = 2u * max_stripe_brows(width / 4);
    for (uint32_t y0 = 0; y0 < height;) {
        const uint32_t rows = std::min(max_px_rows, height - y0);
        hipError_t e = launch_rgb565_baseline(static_cast<const uint8_t*>(d_src) + (size_t)y0 * src_pitch, src_pitch,
                                              static_cast<uint8_t*>(d_dst) + (size_t)y0 * dst_pitch, dst_pitch,
                                              width, rows, (hipStream_t)stream,
                                              y0 == 0 ? (hipEvent_t)start_event : nullptr,
                                              y0 + rows == height ? (hipEvent_t)stop_event : nullptr);
        if (e != hipSuccess) return fail_hip(e, "rgb565 baseline kernel launch");
        y0 += rows;
    }
    return GOOFY_OK;
}

}  // extern "C"
