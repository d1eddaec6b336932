// serve.cpp -- continuous-batching encode server and its load generator.
//
// The reference has no serving layer (its API is one synchronous call per
// image, GoofyTC/goofy_tc.h:1497-1557).  This one keeps textures and results
// in HBM and turns a stream of independent requests into overlapping batch
// launches:
//
//   submit() --> pending queue --> [scheduler thread]
//                                     | admit in arrival order while the
//                                     | in-flight byte budget and the output
//                                     | arena allow; complete expired
//                                     | requests unlaunched
//                                     v
//                 one batch-kernel launch per admitted group, on a free
//                 stream (up to `streams` groups in flight at once)
//                                     |
//                 event polled done --> completion queue --> poll()
//                 release(ticket) frees the output range and budget share
//   cancel(ticket) takes a request out of the pending queue.
//
// Requests that arrive while groups are running are admitted and launched
// immediately on another stream, so their work overlaps on the device with
// the groups already in flight.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "goofy_hip.h"
#include "internal.h"

using namespace goofy_impl;

namespace {

using Clock = std::chrono::steady_clock;

constexpr size_t kAlign = 256;
inline size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// First-fit allocator over one device allocation, with coalescing frees.
class Arena {
public:
    bool init(size_t bytes)
    {
        size_ = align_up(bytes);
        if (hipMalloc(&base_, size_) != hipSuccess) {
            base_ = nullptr;
            return false;
        }
        free_[0] = size_;
        return true;
    }
    ~Arena()
    {
        if (base_) (void)hipFree(base_);
    }
    // returns offset or SIZE_MAX
    size_t alloc(size_t bytes)
    {
        bytes = align_up(std::max<size_t>(bytes, 1));
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                const size_t off = it->first, rest = it->second - bytes;
                free_.erase(it);
                if (rest) free_[off + bytes] = rest;
                return off;
            }
        }
        return SIZE_MAX;
    }
    void release(size_t off, size_t bytes)
    {
        bytes = align_up(std::max<size_t>(bytes, 1));
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + bytes == next->first) {
            bytes += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += bytes;
                return;
            }
        }
        free_[off] = bytes;
    }
    uint8_t* base() const { return static_cast<uint8_t*>(base_); }
    size_t size() const { return size_; }

private:
    void* base_ = nullptr;
    size_t size_ = 0;
    std::map<size_t, size_t> free_;
};

struct Req {
    uint64_t ticket = 0;
    goofy_request r{};
    size_t in_bytes = 0, out_bytes = 0, arena_off = SIZE_MAX;
    double t_arrival = 0, t_submit = 0, t_launch = 0;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    TexDesc* h_desc = nullptr;  // pinned staging
    TexDesc* d_desc = nullptr;
    std::vector<Req> reqs;
    bool busy = false;
};

}  // namespace

struct goofy_server {
    goofy_server_config cfg{};
    LaunchCfg launch{};
    bool auto_shape = true;  // no shape configured: auto_cfg per launch
    int device = 0, num_cus = 0;
    Clock::time_point t0;
    Arena arena;
    std::vector<Slot> slots;

    std::mutex mu;  // everything below up to `stop`
    std::condition_variable cv_work, cv_idle;
    std::deque<Req> pending;  // submitted, not yet admitted, in arrival order
    std::deque<goofy_completion> completions;
    struct Held {
        size_t off, bytes;
        bool done;  // kernel finished: the range may be released
    };
    std::unordered_map<uint64_t, Held> outstanding;  // admitted, unreleased tickets
    uint64_t next_ticket = 1;
    size_t budget_used = 0, peak_used = 0;
    goofy_server_stats st{};
    uint64_t unfinished = 0;     // submitted, not yet completed
    uint64_t inflight_reqs = 0;  // admitted, not yet completed
    double min_deadline = 0;     // earliest deadline among pending requests (0 = none)
    // Nothing runs and the oldest pending request waits for budget or arena
    // space held by completed, unreleased outputs: only releases can help.
    bool stalled = false;

    std::atomic<bool> stop{false};
    std::thread thread;

    double now_us() const { return std::chrono::duration<double, std::micro>(Clock::now() - t0).count(); }
    void run();
    bool launch_group(Slot& s, std::vector<Req>& group);
    void finish(Slot& s, int status);
    void complete_unlaunched(const Req& q, int status, double t);  // caller holds mu
    void expire_pending(double t);                                  // caller holds mu
};

void goofy_server::complete_unlaunched(const Req& q, int status, double t)
{
    goofy_completion c{};
    c.ticket = q.ticket;
    c.user_tag = q.r.user_tag;
    c.status = status;
    c.t_arrival_us = q.t_arrival;
    c.t_submit_us = q.t_submit;
    c.t_done_us = t;
    completions.push_back(c);
    ++st.completed;
    if (status == GOOFY_ERR_CANCELLED) ++st.cancelled;
    else ++st.expired;
    if (--unfinished == 0) cv_idle.notify_all();
}

void goofy_server::expire_pending(double t)
{
    if (min_deadline == 0 || t < min_deadline) return;
    double next = 0;
    for (auto it = pending.begin(); it != pending.end();) {
        const double d = it->r.deadline_us;
        if (d > 0 && d <= t) {
            complete_unlaunched(*it, GOOFY_ERR_DEADLINE, t);
            it = pending.erase(it);
        } else {
            if (d > 0 && (next == 0 || d < next)) next = d;
            ++it;
        }
    }
    min_deadline = next;
}

void goofy_server::finish(Slot& s, int status)
{
    const double t = now_us();
    std::lock_guard<std::mutex> lock(mu);
    for (Req& q : s.reqs) {
        budget_used -= q.in_bytes;  // the input may be reused once encoded
        outstanding[q.ticket].done = true;
        goofy_completion c{};
        c.ticket = q.ticket;
        c.user_tag = q.r.user_tag;
        c.status = status;
        c.dst = arena.base() + q.arena_off;
        c.dst_bytes = q.out_bytes;
        c.t_arrival_us = q.t_arrival;
        c.t_submit_us = q.t_submit;
        c.t_launch_us = q.t_launch;
        c.t_done_us = t;
        completions.push_back(c);
        ++st.completed;
        --unfinished;
        --inflight_reqs;
    }
    s.reqs.clear();
    s.busy = false;
    if (unfinished == 0) cv_idle.notify_all();
}

bool goofy_server::launch_group(Slot& s, std::vector<Req>& group)
{
    const double t = now_us();
    hipError_t e = hipSuccess;
    if (cfg.batching) {
        // Groups of up to kInlineTextures travel in the kernel arguments (no
        // descriptor copy ahead of the kernel); larger ones through the slot's
        // pinned staging and one host-to-device copy.
        const bool inline_list = group.size() <= (size_t)kInlineTextures;
        InlineWorklist wl;
        TexDesc* descs = inline_list ? wl.d : s.h_desc;
        uint32_t* begins = inline_list ? wl.tile_begin : worklist_tile_begin(s.h_desc, group.size());
        LaunchCfg L = launch;
        if (auto_shape) {  // the shape follows the group's size
            uint64_t blocks = 0;
            for (const Req& q : group) blocks += (uint64_t)(q.r.width / 4) * (q.r.height / 4);
            L = auto_cfg(blocks, num_cus);
        }
        const uint64_t tile = (uint64_t)tile_blocks(L);
        uint32_t tiles = 0;
        bool alpha = false;  // the group holds DXT5 / ETC2 RGBA8 requests
        for (size_t i = 0; i < group.size(); ++i) {
            const goofy_request& r = group[i].r;
            TexDesc d{};
            d.src = (uint64_t)(uintptr_t)r.src;
            d.dst = (uint64_t)(uintptr_t)(arena.base() + group[i].arena_off);
            d.pitch = r.pitch;
            d.bw = r.width / 4;
            d.nblocks = d.bw * (r.height / 4);
            d.magic = div_magic(d.bw);
            d.fmt = (uint32_t)r.format;
            alpha |= r.format >= GOOFY_FORMAT_DXT5;
            d.tile_begin = tiles;
            tiles += (uint32_t)((d.nblocks + tile - 1) / tile);
            descs[i] = d;
            begins[i] = d.tile_begin;
        }
        begins[group.size()] = tiles;
        if (inline_list) {
            e = launch_batch_inline(L, wl, (uint32_t)group.size(), tiles, num_cus, s.stream, alpha);
        } else {
            e = hipMemcpyAsync(s.d_desc, s.h_desc, worklist_bytes(group.size()), hipMemcpyHostToDevice, s.stream);
            if (e == hipSuccess) e = launch_batch(L, s.d_desc, (uint32_t)group.size(), tiles, num_cus, s.stream, alpha);
        }
    } else {
        for (Req& q : group) {
            const uint32_t bw = q.r.width / 4, nb = bw * (q.r.height / 4);
            e = launch_plain(q.r.format, auto_shape ? auto_cfg(nb, num_cus) : launch, q.r.src, q.r.pitch,
                             arena.base() + q.arena_off, bw, nb, s.stream);
            if (e != hipSuccess) break;
        }
    }
    if (e == hipSuccess) e = hipEventRecord(s.done, s.stream);
    for (Req& q : group) q.t_launch = t;
    s.reqs.swap(group);
    s.busy = true;
    {
        std::lock_guard<std::mutex> lock(mu);
        st.launches += cfg.batching ? 1 : s.reqs.size();  // kernel launches
    }
    if (e != hipSuccess) {
        // report the failure on every request of the group
        (void)hipStreamSynchronize(s.stream);
        finish(s, GOOFY_ERR_HIP);
        return false;
    }
    return true;
}

void goofy_server::run()
{
    (void)hipSetDevice(device);
    const int max_batch = cfg.batching ? cfg.max_batch_textures : 1 << 30;
    constexpr double kSpinUs = 2000.0;
    double idle_since = -1;
    std::vector<Req> group;
    while (!stop.load(std::memory_order_acquire)) {
        bool progress = false;
        for (Slot& s : slots) {
            if (!s.busy) continue;
            const hipError_t q = hipEventQuery(s.done);
            if (q == hipErrorNotReady) continue;
            // a failed query completes the group with an error instead of
            // leaving its slot busy forever
            finish(s, q == hipSuccess ? GOOFY_OK : GOOFY_ERR_HIP);
            progress = true;
        }
        // admission: arrival order, stop at the first request that does not fit
        while (true) {
            Slot* free_slot = nullptr;
            bool any_busy = false;
            for (Slot& s : slots) {
                if (!s.busy && !free_slot) free_slot = &s;
                any_busy |= s.busy;
            }
            group.clear();
            {
                std::lock_guard<std::mutex> lock(mu);
                const double t = now_us();
                expire_pending(t);
                while (free_slot && !pending.empty() && (int)group.size() < max_batch) {
                    Req& q = pending.front();
                    const size_t need = q.in_bytes + align_up(q.out_bytes);
                    if (budget_used + need > cfg.hbm_budget_bytes) break;
                    const size_t off = arena.alloc(q.out_bytes);
                    if (off == SIZE_MAX) break;
                    q.arena_off = off;
                    budget_used += need;
                    peak_used = std::max(peak_used, budget_used);
                    outstanding[q.ticket] = Held{off, q.out_bytes, false};
                    group.push_back(std::move(q));
                    pending.pop_front();
                }
                st.admitted += group.size();
                inflight_reqs += group.size();
                // nothing runs, yet the head does not fit: the budget or the
                // arena is held by completed outputs that await release
                const bool now_stalled = group.empty() && free_slot && !any_busy && !pending.empty();
                if (now_stalled && !stalled) cv_idle.notify_all();
                stalled = now_stalled;
            }
            if (group.empty()) break;
            launch_group(*free_slot, group);
            progress = true;
        }
        if (progress) {
            idle_since = -1;
            continue;
        }
        bool any_busy = false;
        for (Slot& s : slots) any_busy |= s.busy;
        const double t = now_us();
        if (any_busy || idle_since < 0 || t - idle_since < kSpinUs) {
            // poll events / the pending queue again; spinning a little before
            // sleeping keeps wake-up latency off the request path
            if (!any_busy && idle_since < 0) idle_since = t;
            std::this_thread::yield();
        } else {
            // woken by submits, releases and cancels; the timeout bounds how
            // late an expired deadline is reported
            std::unique_lock<std::mutex> lock(mu);
            cv_work.wait_for(lock, std::chrono::microseconds(500), [&] {
                return (!pending.empty() && !stalled) || stop.load(std::memory_order_acquire);
            });
            idle_since = -1;
        }
    }
    for (Slot& s : slots)
        if (s.busy) finish(s, hipEventSynchronize(s.done) == hipSuccess ? GOOFY_OK : GOOFY_ERR_HIP);
}

extern "C" {

int goofy_server_create(const goofy_server_config* ucfg, goofy_server** out)
{
    clear_error();
    if (!out) return fail(GOOFY_ERR_INVALID, "null argument");
    *out = nullptr;
    goofy_server_config c = ucfg ? *ucfg : goofy_server_config{};
    if (c.hbm_budget_bytes == 0) c.hbm_budget_bytes = 32ull << 30;
    if (c.arena_bytes == 0) c.arena_bytes = c.hbm_budget_bytes / 4;
    if (c.streams <= 0) c.streams = 4;
    if (c.max_batch_textures <= 0) c.max_batch_textures = 256;
    if (c.max_pending <= 0) c.max_pending = 65536;
    if (c.streams > 64 || c.max_batch_textures > 65536) return fail(GOOFY_ERR_INVALID, "streams <= 64, batch <= 65536");
    LaunchCfg lc{1, 8, 0, 1, 0};
    const bool auto_shape = !(c.launch.blocks_per_lane || c.launch.wg_waves || c.launch.waves_per_cu ||
                              c.launch.nontemporal || c.launch.tile_order);
    if (!auto_shape) {
        lc = to_cfg(c.launch);
        if (!cfg_valid(lc)) return fail(GOOFY_ERR_INVALID, "invalid launch config");
    }
    DeviceInfo di;
    int rc = device_info(&di);
    if (rc != GOOFY_OK) return rc;
    goofy_server* s = new goofy_server();
    s->cfg = c;
    s->launch = lc;
    s->auto_shape = auto_shape;
    s->device = di.device;
    s->num_cus = di.num_cus;
    s->t0 = Clock::now();
    s->st.budget_bytes = c.hbm_budget_bytes;
    if (!s->arena.init(c.arena_bytes)) {
        delete s;
        return fail(GOOFY_ERR_OUT_OF_MEMORY, "hipMalloc(output arena) failed");
    }
    s->slots.resize((size_t)c.streams);
    const size_t desc_bytes = worklist_bytes((size_t)c.max_batch_textures);
    for (Slot& sl : s->slots) {
        hipError_t e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
        if (e == hipSuccess) e = hipHostMalloc((void**)&sl.h_desc, desc_bytes, hipHostMallocDefault);
        if (e == hipSuccess) e = hipMalloc((void**)&sl.d_desc, desc_bytes);
        if (e != hipSuccess) {
            rc = fail_hip(e, "server stream/event/descriptor setup");
            goofy_server_destroy(s);
            return rc;
        }
    }
    s->thread = std::thread([s] { s->run(); });
    *out = s;
    return GOOFY_OK;
}

static int validate_request(const goofy_server* s, const goofy_request* r, Req* q)
{
    if (!format_valid(r->format)) return fail(GOOFY_ERR_INVALID, "unknown format");
    if (r->width % 4 != 0) return fail(GOOFY_ERR_WIDTH, "width must be a multiple of 4");
    if (r->height % 4 != 0) return fail(GOOFY_ERR_HEIGHT, "height must be a multiple of 4");
    if (r->width == 0 || r->height == 0) return fail(GOOFY_ERR_INVALID, "empty texture");
    if (!r->src || (uintptr_t)r->src % 16 != 0) return fail(GOOFY_ERR_INVALID, "src must be 16-byte aligned");
    if (r->pitch < (size_t)r->width * 4 || r->pitch % 16 != 0)
        return fail(GOOFY_ERR_INVALID, "pitch must be >= 4*width and a multiple of 16");
    const uint64_t bw = r->width / 4, nb = bw * (r->height / 4);
    if (!magic_ok(bw, nb) || nb > 0xFFFFFFFFull) return fail(GOOFY_ERR_TOO_LARGE, "texture too large");
    q->r = *r;
    q->in_bytes = r->pitch * r->height;
    q->out_bytes = nb * block_bytes(r->format);
    if (q->in_bytes + align_up(q->out_bytes) > s->cfg.hbm_budget_bytes || align_up(q->out_bytes) > s->arena.size())
        return fail(GOOFY_ERR_INVALID, "request larger than the server's budget or arena");
    return GOOFY_OK;
}

int goofy_server_submit_many(goofy_server* s, const goofy_request* reqs, int n, uint64_t* tickets, int* accepted)
{
    clear_error();
    if (!s || (n > 0 && !reqs) || !accepted) return fail(GOOFY_ERR_INVALID, "null argument");
    *accepted = 0;
    int rc = GOOFY_OK;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        const double t = s->now_us();
        for (int i = 0; i < n; ++i) {
            Req q;
            if ((rc = validate_request(s, &reqs[i], &q)) != GOOFY_OK) break;
            if (s->pending.size() >= (size_t)s->cfg.max_pending) {
                ++s->st.rejected;
                rc = fail(GOOFY_ERR_BUSY, "pending queue full");
                break;
            }
            q.ticket = s->next_ticket++;
            q.t_submit = t;
            q.t_arrival = q.r.arrival_us > 0 ? q.r.arrival_us : t;
            const double d = q.r.deadline_us;
            if (d > 0 && (s->min_deadline == 0 || d < s->min_deadline)) s->min_deadline = d;
            if (tickets) tickets[i] = q.ticket;
            s->pending.push_back(q);
            ++s->st.submitted;
            ++s->unfinished;
            ++*accepted;
        }
    }
    if (*accepted) s->cv_work.notify_one();
    return rc;
}

int goofy_server_submit(goofy_server* s, const goofy_request* r, uint64_t* ticket)
{
    if (!r) return fail(GOOFY_ERR_INVALID, "null argument");
    int accepted = 0;
    return goofy_server_submit_many(s, r, 1, ticket, &accepted);
}

int goofy_server_poll(goofy_server* s, goofy_completion* out, int max, int* count)
{
    clear_error();
    if (!s || !count || (max > 0 && !out)) return fail(GOOFY_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lock(s->mu);
    int n = 0;
    while (n < max && !s->completions.empty()) {
        out[n++] = s->completions.front();
        s->completions.pop_front();
    }
    *count = n;
    return GOOFY_OK;
}

// caller holds s->mu
static int release_locked(goofy_server* s, uint64_t ticket)
{
    auto it = s->outstanding.find(ticket);
    if (it == s->outstanding.end()) return fail(GOOFY_ERR_INVALID, "unknown or already released ticket");
    if (!it->second.done) return fail(GOOFY_ERR_BUSY, "request not completed yet");
    s->arena.release(it->second.off, it->second.bytes);
    s->budget_used -= align_up(it->second.bytes);
    s->outstanding.erase(it);
    ++s->st.released;
    s->stalled = false;  // the scheduler looks at the pending queue again
    return GOOFY_OK;
}

int goofy_server_release(goofy_server* s, uint64_t ticket)
{
    clear_error();
    if (!s) return fail(GOOFY_ERR_INVALID, "null server");
    int rc;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        rc = release_locked(s, ticket);
    }
    if (rc == GOOFY_OK) s->cv_work.notify_one();
    return rc;
}

int goofy_server_release_many(goofy_server* s, const uint64_t* tickets, int n)
{
    clear_error();
    if (!s || (n > 0 && !tickets)) return fail(GOOFY_ERR_INVALID, "null argument");
    int rc = GOOFY_OK;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        for (int i = 0; i < n; ++i) {
            const int r = release_locked(s, tickets[i]);
            if (r != GOOFY_OK && rc == GOOFY_OK) rc = r;
        }
    }
    if (n > 0) s->cv_work.notify_one();
    return rc;
}

int goofy_server_cancel(goofy_server* s, uint64_t ticket)
{
    clear_error();
    if (!s) return fail(GOOFY_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> lock(s->mu);
    // newest first: a caller usually cancels what it submitted last
    for (auto it = s->pending.rbegin(); it != s->pending.rend(); ++it) {
        if (it->ticket != ticket) continue;
        s->complete_unlaunched(*it, GOOFY_ERR_CANCELLED, s->now_us());
        s->pending.erase(std::next(it).base());
        s->stalled = false;  // a new head may fit
        s->cv_work.notify_one();
        return GOOFY_OK;
    }
    if (s->outstanding.count(ticket)) return fail(GOOFY_ERR_BUSY, "request already launched");
    return fail(GOOFY_ERR_INVALID, "unknown, completed or already cancelled ticket");
}

int goofy_server_wait_idle(goofy_server* s, int timeout_ms)
{
    clear_error();
    if (!s) return fail(GOOFY_ERR_INVALID, "null server");
    std::unique_lock<std::mutex> lock(s->mu);
    auto done_or_stuck = [&] { return s->unfinished == 0 || s->stalled; };
    if (timeout_ms < 0) s->cv_idle.wait(lock, done_or_stuck);
    else if (!s->cv_idle.wait_for(lock, std::chrono::milliseconds(timeout_ms), done_or_stuck))
        return fail(GOOFY_ERR_BUSY, "requests still in flight");
    if (s->unfinished != 0)
        return fail(GOOFY_ERR_BUSY, "pending requests wait for completed outputs to be released");
    return GOOFY_OK;
}

int goofy_server_stats_get(goofy_server* s, goofy_server_stats* out)
{
    clear_error();
    if (!s || !out) return fail(GOOFY_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lock(s->mu);
    *out = s->st;
    out->pending = s->pending.size();
    out->inflight_requests = s->inflight_reqs;
    out->inflight_bytes = s->budget_used;
    out->peak_inflight_bytes = s->peak_used;
    return GOOFY_OK;
}

double goofy_server_now_us(const goofy_server* s) { return s ? s->now_us() : 0.0; }

int goofy_server_destroy(goofy_server* s)
{
    clear_error();
    if (!s) return GOOFY_OK;
    if (s->thread.joinable()) {
        s->stop.store(true, std::memory_order_release);
        s->cv_work.notify_all();
        s->thread.join();
    }
    for (Slot& sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.h_desc) (void)hipHostFree(sl.h_desc);
        if (sl.d_desc) (void)hipFree(sl.d_desc);
    }
    delete s;
    return GOOFY_OK;
}

// ---- load generator --------------------------------------------------------------------

int goofy_serve_simulate(const goofy_serve_sim_config* cfg, goofy_serve_sim_result* res)
{
    clear_error();
    if (!cfg || !res || !cfg->pool || cfg->pool_count <= 0 || cfg->duration_s <= 0)
        return fail(GOOFY_ERR_INVALID, "bad simulation config");
    if (cfg->rate_per_s <= 0 && cfg->closed_loop_depth <= 0)
        return fail(GOOFY_ERR_INVALID, "need a rate or a closed-loop depth");
    *res = goofy_serve_sim_result{};
    goofy_server* srv = nullptr;
    int rc = goofy_server_create(&cfg->server, &srv);
    if (rc != GOOFY_OK) return rc;

    // arrivals: exponential gaps (Poisson process), uniform pool choice
    uint64_t sm = cfg->seed * 0x9e3779b97f4a7c15ull + 1;
    auto next_u64 = [&]() {
        uint64_t z = (sm += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    };
    auto next_unit = [&]() { return ((next_u64() >> 11) + 0.5) * (1.0 / 9007199254740992.0); };

    std::vector<double> lat, queue, lag, lat_submit;
    std::vector<uint8_t> host;
    std::atomic<bool> gen_done{false};
    std::atomic<uint64_t> outstanding{0};
    uint64_t verified = 0, mismatches = 0, completed = 0;
    double mpix = 0;
    hipStream_t vstream = nullptr;
    (void)hipStreamCreateWithFlags(&vstream, hipStreamNonBlocking);

    // completion consumer: latency bookkeeping, sampled verification, release
    std::thread consumer([&] {
        std::vector<goofy_completion> buf(1024);
        std::vector<uint64_t> done;
        uint64_t seen = 0;
        while (true) {
            int n = 0;
            goofy_server_poll(srv, buf.data(), (int)buf.size(), &n);
            done.clear();
            for (int i = 0; i < n; ++i) {
                const goofy_completion& c = buf[i];
                lat.push_back(c.t_done_us - c.t_arrival_us);
                queue.push_back(c.t_launch_us - c.t_arrival_us);
                lag.push_back(c.t_submit_us - c.t_arrival_us);
                lat_submit.push_back(c.t_done_us - c.t_submit_us);
                const goofy_texture_desc& t = cfg->pool[c.user_tag];
                mpix += (double)t.width * t.height / 1e6;
                if (c.status != GOOFY_OK) ++mismatches;
                else if (cfg->pool_expected && cfg->verify_every > 0 && seen % (uint64_t)cfg->verify_every == 0) {
                    host.resize(c.dst_bytes);
                    if (hipMemcpyAsync(host.data(), c.dst, c.dst_bytes, hipMemcpyDeviceToHost, vstream) == hipSuccess &&
                        hipStreamSynchronize(vstream) == hipSuccess) {
                        ++verified;
                        if (std::memcmp(host.data(), cfg->pool_expected[c.user_tag], c.dst_bytes) != 0) ++mismatches;
                    } else {
                        ++mismatches;
                    }
                }
                ++seen;
                ++completed;
                done.push_back(c.ticket);
            }
            if (n) {
                goofy_server_release_many(srv, done.data(), (int)done.size());
                outstanding.fetch_sub((uint64_t)n);
            }
            if (n == 0) {
                if (gen_done.load() && outstanding.load() == 0) break;
                std::this_thread::yield();
            }
        }
    });

    uint64_t submitted = 0, rejected = 0;
    const double t_start = goofy_server_now_us(srv);
    const double t_end = t_start + cfg->duration_s * 1e6;
    double next_arrival = t_start;
    double offered_px = 0;
    std::vector<goofy_request> chunk;
    std::vector<uint64_t> tickets;
    while (true) {
        const double now = goofy_server_now_us(srv);
        if (now >= t_end) break;
        chunk.clear();
        if (cfg->rate_per_s > 0) {
            // every arrival that is due by now, in one call, each stamped with
            // its scheduled arrival time
            while (next_arrival <= now && chunk.size() < 4096) {
                const int k = (int)(next_u64() % (uint64_t)cfg->pool_count);
                const goofy_texture_desc& t = cfg->pool[k];
                chunk.push_back(goofy_request{t.src, t.width, t.height, t.pitch, t.format, (uint64_t)k, next_arrival,
                                              0.0});
                next_arrival += -std::log(next_unit()) / cfg->rate_per_s * 1e6;
            }
            if (chunk.empty()) {
                // sleep only through long gaps: a sleep can overrun by the
                // timer slack (about 50 us), shorter waits spin
                const double gap = next_arrival - now;
                if (gap > 300.0) std::this_thread::sleep_for(std::chrono::microseconds((int64_t)(gap - 200.0)));
                else std::this_thread::yield();
                continue;
            }
        } else {
            const uint64_t out = outstanding.load();
            if (out >= (uint64_t)cfg->closed_loop_depth) {
                std::this_thread::yield();
                continue;
            }
            const uint64_t room = std::min<uint64_t>((uint64_t)cfg->closed_loop_depth - out, 256);
            for (uint64_t j = 0; j < room; ++j) {
                const int k = (int)(next_u64() % (uint64_t)cfg->pool_count);
                const goofy_texture_desc& t = cfg->pool[k];
                chunk.push_back(goofy_request{t.src, t.width, t.height, t.pitch, t.format, (uint64_t)k, 0.0, 0.0});
            }
        }
        tickets.resize(chunk.size());
        outstanding.fetch_add(chunk.size());
        int accepted = 0;
        const int s = goofy_server_submit_many(srv, chunk.data(), (int)chunk.size(), tickets.data(), &accepted);
        submitted += (uint64_t)accepted;
        for (int j = 0; j < accepted; ++j) offered_px += (double)chunk[j].width * chunk[j].height;
        if (accepted < (int)chunk.size()) {
            outstanding.fetch_sub(chunk.size() - (size_t)accepted);
            rejected += chunk.size() - (size_t)accepted;
            if (s != GOOFY_ERR_BUSY) {
                rc = s;
                break;
            }
        }
    }
    const double t_gen_end = goofy_server_now_us(srv);
    gen_done.store(true);
    consumer.join();
    const double t_all_done = goofy_server_now_us(srv);
    goofy_server_stats st{};
    goofy_server_stats_get(srv, &st);
    goofy_server_destroy(srv);
    if (vstream) (void)hipStreamDestroy(vstream);
    if (rc != GOOFY_OK) return rc;

    std::vector<double> sl = lat, sq = queue;
    std::sort(sl.begin(), sl.end());
    std::sort(sq.begin(), sq.end());
    std::sort(lag.begin(), lag.end());
    std::sort(lat_submit.begin(), lat_submit.end());
    auto pct = [](const std::vector<double>& v, double p) {
        if (v.empty()) return 0.0;
        const size_t i = std::min(v.size() - 1, (size_t)std::floor(p * (double)(v.size() - 1) + 0.5));
        return v[i];
    };
    double sum = 0;
    for (double v : sl) sum += v;
    res->submitted = submitted;
    res->completed = completed;
    res->rejected = rejected;
    res->verified = verified;
    res->mismatches = mismatches;
    res->launches = st.launches;
    res->wall_s = (t_all_done - t_start) / 1e6;
    res->megapixels = mpix;
    res->mps = mpix / res->wall_s;
    res->offered_mps = offered_px / 1e6 / ((t_gen_end - t_start) / 1e6);
    res->lat_mean_us = sl.empty() ? 0 : sum / (double)sl.size();
    res->lat_p50_us = pct(sl, 0.50);
    res->lat_p90_us = pct(sl, 0.90);
    res->lat_p99_us = pct(sl, 0.99);
    res->lat_max_us = sl.empty() ? 0 : sl.back();
    res->queue_p50_us = pct(sq, 0.50);
    res->queue_p99_us = pct(sq, 0.99);
    res->mean_textures_per_launch = st.launches ? (double)st.admitted / (double)st.launches : 0;
    res->peak_inflight_bytes = st.peak_inflight_bytes;
    res->gen_lag_p50_us = pct(lag, 0.50);
    res->gen_lag_p99_us = pct(lag, 0.99);
    res->gen_lag_max_us = lag.empty() ? 0 : lag.back();
    res->submit_lat_p50_us = pct(lat_submit, 0.50);
    res->submit_lat_p99_us = pct(lat_submit, 0.99);
    return GOOFY_OK;
}

}  // extern "C"
