// serve_fake_backend.cpp -- a fake device for CPU tests of the serving
// scheduler (goofy_amd/csrc/serve.cpp): the HIP runtime calls the scheduler
// makes (declared in fake_hip/hip/hip_runtime_api.h) and the library's
// internal helpers it links against (internal.h: errors, device info, launch
// shapes and the three kernel launchers).
//
// Streams are worker threads running their operations in order, so kernels,
// copies and event records on different streams overlap as on a GPU, and
// ThreadSanitizer sees the scheduler's real concurrency.  The "encode kernel"
// writes, for every 4x4 block, a 64-bit FNV-1a digest of the block's 16 pixels
// and its format (twice for the 16-byte formats): any mix-up of sources,
// pitches, formats or output ranges shows as a wrong digest.  Failure
// injection (allocation failure, launch failure, a kernel that faults its
// stream) and a kernel delay come from FakeHipControl.
#include <hip/hip_runtime_api.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <mutex>
#include <string>
#include <thread>

#include "internal.h"

namespace {

std::mutex g_ctl_mu;
FakeHipControl g_ctl{0, -1, -1, -1};
std::atomic<uint64_t> g_kernels_run{0};

// Counts down a control field; true when it reaches the failing call.
bool take_failure(int FakeHipControl::*field)
{
    std::lock_guard<std::mutex> lock(g_ctl_mu);
    int& n = g_ctl.*field;
    if (n < 0) return false;
    if (n == 0) {
        n = -1;
        return true;
    }
    --n;
    return false;
}

int kernel_delay_us()
{
    std::lock_guard<std::mutex> lock(g_ctl_mu);
    return g_ctl.kernel_delay_us;
}

}  // namespace

struct FakeStream {
    std::mutex mu;
    std::condition_variable cv, cv_idle;
    std::deque<std::function<void()>> ops;
    bool stop = false, running = false;
    bool faulted = false;  // a kernel on this stream faulted (worker thread only)
    std::thread worker;

    FakeStream() : worker([this] { loop(); }) {}
    ~FakeStream()
    {
        {
            std::lock_guard<std::mutex> lock(mu);
            stop = true;
        }
        cv.notify_all();
        worker.join();
    }
    void push(std::function<void()> op)
    {
        {
            std::lock_guard<std::mutex> lock(mu);
            ops.push_back(std::move(op));
        }
        cv.notify_all();
    }
    void drain()
    {
        std::unique_lock<std::mutex> lock(mu);
        cv_idle.wait(lock, [&] { return ops.empty() && !running; });
    }
    void loop()
    {
        std::unique_lock<std::mutex> lock(mu);
        while (true) {
            cv.wait(lock, [&] { return stop || !ops.empty(); });
            if (ops.empty()) return;  // stop, nothing left
            std::function<void()> op = std::move(ops.front());
            ops.pop_front();
            running = true;
            lock.unlock();
            op();
            lock.lock();
            running = false;
            if (ops.empty()) cv_idle.notify_all();
        }
    }
};

struct FakeEvent {
    std::mutex mu;
    std::condition_variable cv;
    uint64_t recorded = 0, reached = 0;
    bool error = false;
};

namespace {

FakeStream* null_stream()
{
    static FakeStream* s = new FakeStream();  // lives until exit
    return s;
}

FakeStream* S(hipStream_t s) { return s ? s : null_stream(); }

uint64_t block_digest(const uint8_t* src, uint64_t pitch, uint32_t bw, uint64_t i, uint32_t fmt)
{
    const uint64_t by = i / bw, bx = i % bw;
    uint64_t h = 1469598103934665603ull ^ fmt;
    for (int r = 0; r < 4; ++r) {
        const uint8_t* row = src + (by * 4 + r) * pitch + bx * 16;
        for (int k = 0; k < 16; ++k) h = (h ^ row[k]) * 1099511628211ull;
    }
    return h;
}

}  // namespace

// The fake "encoder" of one texture: the digest of each block, in the block
// layout of the real formats (8 or 16 bytes per block, row-major).
static void fake_encode(uint32_t fmt, const void* src, uint64_t pitch, void* dst, uint32_t bw, uint64_t nblocks)
{
    const size_t bb = goofy_impl::block_bytes((int)fmt);
    for (uint64_t i = 0; i < nblocks; ++i) {
        const uint64_t h = block_digest(static_cast<const uint8_t*>(src), pitch, bw, i, fmt);
        uint8_t* out = static_cast<uint8_t*>(dst) + i * bb;
        std::memcpy(out, &h, 8);
        if (bb == 16) {
            const uint64_t h2 = ~h;
            std::memcpy(out + 8, &h2, 8);
        }
    }
}

// The kernel body that runs on a stream: delay, fault injection, then work.
static void run_kernel(FakeStream* st, const std::function<void()>& work)
{
    const int d = kernel_delay_us();
    if (d > 0) std::this_thread::sleep_for(std::chrono::microseconds(d));
    if (take_failure(&FakeHipControl::fault_kernel_after)) {
        st->faulted = true;
        return;
    }
    if (!st->faulted) work();
    ++g_kernels_run;
}

// ---- the HIP runtime slice ------------------------------------------------------------

hipError_t hipMalloc(void** ptr, size_t bytes)
{
    if (!ptr) return hipErrorInvalidValue;
    if (take_failure(&FakeHipControl::fail_malloc_after)) return hipErrorOutOfMemory;
    *ptr = std::aligned_alloc(256, (bytes + 255) / 256 * 256 + 256);
    return *ptr ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void* ptr)
{
    std::free(ptr);
    return hipSuccess;
}
hipError_t hipHostMalloc(void** ptr, size_t bytes, unsigned int)
{
    if (!ptr) return hipErrorInvalidValue;
    *ptr = std::aligned_alloc(256, (bytes + 255) / 256 * 256 + 256);
    return *ptr ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipHostFree(void* ptr)
{
    std::free(ptr);
    return hipSuccess;
}
hipError_t hipSetDevice(int device) { return device == 0 ? hipSuccess : hipErrorInvalidDevice; }
hipError_t hipStreamCreateWithFlags(hipStream_t* stream, unsigned int)
{
    *stream = new FakeStream();
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t stream)
{
    if (stream) {
        stream->drain();
        delete stream;
    }
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t stream)
{
    S(stream)->drain();
    return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* event, unsigned int)
{
    *event = new FakeEvent();
    return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t event)
{
    delete event;
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t stream)
{
    FakeStream* st = S(stream);
    uint64_t gen;
    {
        std::lock_guard<std::mutex> lock(e->mu);
        gen = ++e->recorded;
    }
    st->push([e, st, gen] {
        std::lock_guard<std::mutex> lock(e->mu);
        if (gen > e->reached) e->reached = gen;
        e->error = st->faulted;
        e->cv.notify_all();
    });
    return hipSuccess;
}
hipError_t hipEventQuery(hipEvent_t e)
{
    std::lock_guard<std::mutex> lock(e->mu);
    if (e->reached < e->recorded) return hipErrorNotReady;
    return e->error ? hipErrorLaunchFailure : hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t e)
{
    std::unique_lock<std::mutex> lock(e->mu);
    const uint64_t want = e->recorded;
    e->cv.wait(lock, [&] { return e->reached >= want; });
    return e->error ? hipErrorLaunchFailure : hipSuccess;
}
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind, hipStream_t stream)
{
    S(stream)->push([dst, src, bytes] { std::memcpy(dst, src, bytes); });
    return hipSuccess;
}
const char* hipGetErrorString(hipError_t e)
{
    switch (e) {
    case hipSuccess: return "hipSuccess";
    case hipErrorOutOfMemory: return "hipErrorOutOfMemory (fake)";
    case hipErrorLaunchFailure: return "hipErrorLaunchFailure (fake)";
    case hipErrorNotReady: return "hipErrorNotReady";
    default: return "fake HIP error";
    }
}

void fake_hip_set(const FakeHipControl& c)
{
    std::lock_guard<std::mutex> lock(g_ctl_mu);
    g_ctl = c;
}
FakeHipControl fake_hip_get()
{
    std::lock_guard<std::mutex> lock(g_ctl_mu);
    return g_ctl;
}
uint64_t fake_hip_kernels_run() { return g_kernels_run.load(); }

// ---- the library internals the scheduler links against -------------------------------

namespace goofy_impl {

thread_local std::string t_last_error;

int fail(int code, const std::string& msg)
{
    t_last_error = msg;
    return code;
}
int fail_hip(hipError_t e, const char* what)
{
    return fail(e == hipErrorOutOfMemory ? GOOFY_ERR_OUT_OF_MEMORY : GOOFY_ERR_HIP,
                std::string(what) + ": " + hipGetErrorString(e));
}
void clear_error() { t_last_error.clear(); }

int device_info(DeviceInfo* out)
{
    out->device = 0;
    out->num_cus = 256;
    out->total_mem = 1ull << 34;
    std::strcpy(out->arch, "fake");
    return GOOFY_OK;
}

LaunchCfg auto_cfg(uint64_t, int) { return LaunchCfg{1, 8, 0, 1, 0}; }

bool cfg_valid(const LaunchCfg& c)
{
    return (c.blocks_per_lane == 1 || c.blocks_per_lane == 2 || c.blocks_per_lane == 4) &&
           (c.wg_waves == 1 || c.wg_waves == 2 || c.wg_waves == 4 || c.wg_waves == 8) && c.waves_per_cu >= 0 &&
           (c.nontemporal == 0 || c.nontemporal == 1) && (c.tile_order == 0 || c.tile_order == 1);
}

// A work-list's textures, read when the kernel runs (as the GPU reads them),
// checked against the tile prefix sums the launch was given.
static void run_worklist(const TexDesc* d, const uint32_t* begins, uint32_t ntex, uint32_t ntiles, int tile)
{
    uint32_t tiles = 0;
    for (uint32_t t = 0; t < ntex; ++t) {
        if (begins[t] != tiles || d[t].tile_begin != tiles) std::abort();  // inconsistent work-list
        tiles += (uint32_t)((d[t].nblocks + (uint64_t)tile - 1) / (uint64_t)tile);
        fake_encode(d[t].fmt, reinterpret_cast<const void*>(d[t].src), d[t].pitch, reinterpret_cast<void*>(d[t].dst),
                    d[t].bw, d[t].nblocks);
    }
    if (begins[ntex] != tiles || tiles != ntiles) std::abort();
}

hipError_t launch_batch(const LaunchCfg& cfg, const TexDesc* d_descs, uint32_t ntex, uint32_t ntiles, int,
                        hipStream_t stream, bool)
{
    if (take_failure(&FakeHipControl::fail_launch_after)) return hipErrorLaunchFailure;
    FakeStream* st = S(stream);
    const int tile = tile_blocks(cfg);
    st->push([st, d_descs, ntex, ntiles, tile] {
        run_kernel(st, [&] { run_worklist(d_descs, worklist_tile_begin(d_descs, ntex), ntex, ntiles, tile); });
    });
    return hipSuccess;
}

hipError_t launch_batch_inline(const LaunchCfg& cfg, const InlineWorklist& wl, uint32_t ntex, uint32_t ntiles, int,
                               hipStream_t stream, bool)
{
    if (take_failure(&FakeHipControl::fail_launch_after)) return hipErrorLaunchFailure;
    FakeStream* st = S(stream);
    const int tile = tile_blocks(cfg);
    st->push([st, wl, ntex, ntiles, tile] {  // the arguments travel by value, as kernel arguments do
        run_kernel(st, [&] { run_worklist(wl.d, wl.tile_begin, ntex, ntiles, tile); });
    });
    return hipSuccess;
}

hipError_t launch_plain(int fmt, const LaunchCfg&, const void* src, size_t pitch, void* dst, uint32_t bw,
                        uint32_t nblocks, hipStream_t stream, hipEvent_t, hipEvent_t)
{
    if (take_failure(&FakeHipControl::fail_launch_after)) return hipErrorLaunchFailure;
    FakeStream* st = S(stream);
    st->push([st, fmt, src, pitch, dst, bw, nblocks] {
        run_kernel(st, [&] { fake_encode((uint32_t)fmt, src, pitch, dst, bw, nblocks); });
    });
    return hipSuccess;
}

}  // namespace goofy_impl

extern "C" const char* goofy_hip_last_error(void) { return goofy_impl::t_last_error.c_str(); }

// The reference digest of a texture's blocks (what the fake device writes).
void fake_expected_blocks(int fmt, const void* src, size_t pitch, unsigned width, unsigned height, void* dst)
{
    fake_encode((uint32_t)fmt, src, pitch, dst, width / 4, (uint64_t)(width / 4) * (height / 4));
}
