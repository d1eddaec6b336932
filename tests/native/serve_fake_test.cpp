// serve_fake_test.cpp -- CPU tests of the serving scheduler (serve.cpp) on the
// fake device of serve_fake_backend.cpp, built under ThreadSanitizer and under
// AddressSanitizer by tests/test_serving_cpu.py.  Each case prints its name
// first; the run ends with "serve fake ok".
#include <hip/hip_runtime_api.h>

#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <future>
#include <map>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "goofy_hip.h"

void fake_expected_blocks(int fmt, const void* src, size_t pitch, unsigned width, unsigned height, void* dst);

#define CHECK(c)                                                                                    \
    do {                                                                                            \
        if (!(c)) {                                                                                 \
            std::printf("FAILED %s:%d: %s (%s)\n", __FILE__, __LINE__, #c, goofy_hip_last_error()); \
            std::fflush(stdout);                                                                    \
            std::exit(1);                                                                           \
        }                                                                                           \
    } while (0)

namespace {

struct Tex {
    std::vector<uint8_t> storage;
    const uint8_t* px = nullptr;
    unsigned w = 0, h = 0;
    size_t pitch = 0;
    int fmt = 0;
    std::vector<uint8_t> expected;
    size_t in_bytes() const { return pitch * h; }
    size_t out_bytes() const { return expected.size(); }
};

std::vector<Tex> make_textures(int n, uint32_t seed, unsigned max_side = 128)
{
    std::mt19937 rng(seed);
    std::vector<Tex> v(n);
    for (int i = 0; i < n; ++i) {
        Tex& t = v[i];
        t.w = 4 * (1 + rng() % (max_side / 4));
        t.h = 4 * (1 + rng() % (max_side / 4));
        t.pitch = (t.w * 4 + 15) / 16 * 16 + 16 * (rng() % 3);
        t.fmt = (int)(rng() % 4);
        t.storage.resize(t.pitch * t.h + 16);
        uint8_t* p = t.storage.data() + ((16 - (uintptr_t)t.storage.data() % 16) % 16);
        for (size_t k = 0; k < t.pitch * t.h; ++k) p[k] = (uint8_t)rng();
        t.px = p;
        t.expected.resize((size_t)(t.w / 4) * (t.h / 4) * (t.fmt >= 2 ? 16 : 8));
        fake_expected_blocks(t.fmt, t.px, t.pitch, t.w, t.h, t.expected.data());
    }
    return v;
}

goofy_request req(const Tex& t, uint64_t tag, double arrival = 0, double deadline = 0)
{
    return goofy_request{t.px, t.w, t.h, t.pitch, t.fmt, tag, arrival, deadline};
}

goofy_server* make_server(size_t budget, size_t arena, int streams, int max_batch, int batching, int max_pending = 0)
{
    goofy_server_config c{};
    c.hbm_budget_bytes = budget;
    c.arena_bytes = arena;
    c.streams = streams;
    c.max_batch_textures = max_batch;
    c.batching = batching;
    c.max_pending = max_pending;
    goofy_server* s = nullptr;
    CHECK(goofy_server_create(&c, &s) == GOOFY_OK && s);
    return s;
}

void set_fake(int delay_us, int fail_malloc = -1, int fail_launch = -1, int fault_kernel = -1)
{
    fake_hip_set(FakeHipControl{delay_us, fail_malloc, fail_launch, fault_kernel});
}

// Polls until `want` completions arrived (or 20 s passed), checking each OK
// completion's blocks against the texture it names; returns them by ticket.
std::map<uint64_t, goofy_completion> collect(goofy_server* s, const std::vector<Tex>& tex, size_t want,
                                             bool release = true)
{
    std::map<uint64_t, goofy_completion> got;
    const auto t_end = std::chrono::steady_clock::now() + std::chrono::seconds(20);
    goofy_completion buf[64];
    while (got.size() < want) {
        CHECK(std::chrono::steady_clock::now() < t_end);
        int n = 0;
        CHECK(goofy_server_poll(s, buf, 64, &n) == GOOFY_OK);
        for (int i = 0; i < n; ++i) {
            const goofy_completion& c = buf[i];
            CHECK(got.count(c.ticket) == 0);
            const Tex& t = tex[c.user_tag % tex.size()];
            CHECK(c.t_arrival_us <= c.t_submit_us + 1e-9 || c.t_arrival_us > 0);
            CHECK(c.t_done_us >= c.t_submit_us);
            if (c.status == GOOFY_OK) {
                CHECK(c.t_launch_us >= c.t_submit_us && c.t_done_us >= c.t_launch_us);
                CHECK(c.dst && c.dst_bytes == t.out_bytes());
                CHECK(std::memcmp(c.dst, t.expected.data(), t.out_bytes()) == 0);
            }
            if (release && (c.status == GOOFY_OK || c.status == GOOFY_ERR_HIP))
                CHECK(goofy_server_release(s, c.ticket) == GOOFY_OK);
            got[c.ticket] = c;
        }
        if (n == 0) std::this_thread::sleep_for(std::chrono::microseconds(100));
    }
    return got;
}

void case_concurrent_clients(int batching)
{
    std::printf("case: 4 client threads, 1200 requests, batching=%d\n", batching);
    std::fflush(stdout);
    set_fake(200);
    const std::vector<Tex> tex = make_textures(64, 1 + batching);
    goofy_server* s = make_server(64ull << 20, 16ull << 20, 4, 8, batching);
    constexpr int kClients = 4, kPer = 300;
    std::atomic<int> submitted{0};
    std::vector<std::thread> clients;
    for (int c = 0; c < kClients; ++c)
        clients.emplace_back([&, c] {
            std::mt19937 rng(100 + c);
            for (int i = 0; i < kPer; ++i) {
                const uint64_t tag = rng() % tex.size();
                // half the requests carry an arrival stamp a little in the past
                const double arrival = (i % 2) ? goofy_server_now_us(s) - 5.0 : 0.0;
                goofy_request r = req(tex[tag], tag, arrival > 0 ? arrival : 0.0);
                uint64_t ticket = 0;
                CHECK(goofy_server_submit(s, &r, &ticket) == GOOFY_OK && ticket > 0);
                ++submitted;
                if (rng() % 8 == 0) std::this_thread::sleep_for(std::chrono::microseconds(rng() % 300));
            }
        });
    const auto got = collect(s, tex, kClients * kPer);
    for (auto& t : clients) t.join();
    CHECK(goofy_server_wait_idle(s, -1) == GOOFY_OK);
    for (const auto& kv : got) {
        CHECK(kv.second.status == GOOFY_OK);
        CHECK(kv.second.t_arrival_us <= kv.second.t_submit_us);
    }
    goofy_server_stats st{};
    CHECK(goofy_server_stats_get(s, &st) == GOOFY_OK);
    CHECK(st.submitted == kClients * kPer && st.completed == st.submitted && st.released == st.submitted);
    CHECK(st.pending == 0 && st.inflight_requests == 0 && st.inflight_bytes == 0);
    CHECK(st.peak_inflight_bytes <= st.budget_bytes);
    CHECK(batching ? st.launches <= st.admitted : st.launches == st.admitted);
    CHECK(goofy_server_destroy(s) == GOOFY_OK);
}

void case_budget_and_queue_limit()
{
    std::printf("case: budget of 3 requests, pending queue of 40, burst of 200\n");
    std::fflush(stdout);
    set_fake(100);
    const std::vector<Tex> tex = make_textures(32, 7);
    size_t biggest = 0;
    for (const Tex& t : tex) biggest = std::max(biggest, t.in_bytes() + (t.out_bytes() + 255) / 256 * 256);
    goofy_server* s = make_server(3 * biggest, 2 * biggest, 3, 4, 1, 40);
    std::vector<goofy_request> rs;
    for (int i = 0; i < 200; ++i) rs.push_back(req(tex[i % tex.size()], (uint64_t)i));
    std::map<uint64_t, goofy_completion> got;
    int next = 0, busy_seen = 0;
    const auto t_end = std::chrono::steady_clock::now() + std::chrono::seconds(30);
    while (got.size() < rs.size()) {
        CHECK(std::chrono::steady_clock::now() < t_end);
        if (next < (int)rs.size()) {
            int accepted = 0;
            const int rc = goofy_server_submit_many(s, rs.data() + next, (int)rs.size() - next, nullptr, &accepted);
            CHECK(rc == GOOFY_OK || rc == GOOFY_ERR_BUSY);
            busy_seen += rc == GOOFY_ERR_BUSY;
            next += accepted;
        }
        goofy_server_stats st{};
        CHECK(goofy_server_stats_get(s, &st) == GOOFY_OK);
        CHECK(st.pending <= 40 && st.inflight_bytes <= st.budget_bytes);
        goofy_completion buf[16];
        int n = 0;
        CHECK(goofy_server_poll(s, buf, 16, &n) == GOOFY_OK);
        for (int i = 0; i < n; ++i) {
            CHECK(buf[i].status == GOOFY_OK);
            const Tex& t = tex[buf[i].user_tag % tex.size()];
            CHECK(std::memcmp(buf[i].dst, t.expected.data(), t.out_bytes()) == 0);
            CHECK(goofy_server_release(s, buf[i].ticket) == GOOFY_OK);
            got[buf[i].ticket] = buf[i];
        }
        if (n == 0) std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
    CHECK(busy_seen > 0);  // the queue limit pushed back
    goofy_server_stats st{};
    CHECK(goofy_server_stats_get(s, &st) == GOOFY_OK);
    CHECK(st.rejected > 0 && st.peak_inflight_bytes <= st.budget_bytes && st.inflight_bytes == 0);
    CHECK(goofy_server_destroy(s) == GOOFY_OK);
}

void case_cancel()
{
    std::printf("case: cancel pending, launched, unknown and cancelled tickets\n");
    std::fflush(stdout);
    set_fake(30000);  // 30 ms kernels: the later requests wait behind the first
    const std::vector<Tex> tex = make_textures(8, 11);
    goofy_server* s = make_server(64ull << 20, 16ull << 20, 1, 1, 1);
    uint64_t t[5];
    for (int i = 0; i < 5; ++i) {
        goofy_request r = req(tex[i], (uint64_t)i);
        CHECK(goofy_server_submit(s, &r, &t[i]) == GOOFY_OK);
    }
    // wait until the first request is launched
    for (int k = 0; k < 2000; ++k) {
        goofy_server_stats st{};
        goofy_server_stats_get(s, &st);
        if (st.admitted >= 1) break;
        std::this_thread::sleep_for(std::chrono::microseconds(500));
    }
    CHECK(goofy_server_cancel(s, t[3]) == GOOFY_OK);
    CHECK(goofy_server_cancel(s, t[0]) == GOOFY_ERR_BUSY);     // running
    CHECK(goofy_server_cancel(s, t[3]) == GOOFY_ERR_INVALID);  // already cancelled
    CHECK(goofy_server_cancel(s, 999999) == GOOFY_ERR_INVALID);
    CHECK(goofy_server_release(s, t[3]) == GOOFY_ERR_INVALID);  // a cancelled request holds no output
    const auto got = collect(s, tex, 5);
    CHECK(got.at(t[3]).status == GOOFY_ERR_CANCELLED && got.at(t[3]).dst == nullptr && got.at(t[3]).t_launch_us == 0);
    for (int i : {0, 1, 2, 4}) CHECK(got.at(t[i]).status == GOOFY_OK);
    CHECK(goofy_server_wait_idle(s, -1) == GOOFY_OK);
    goofy_server_stats st{};
    CHECK(goofy_server_stats_get(s, &st) == GOOFY_OK);
    CHECK(st.cancelled == 1 && st.completed == 5 && st.admitted == 4 && st.released == 4 && st.inflight_bytes == 0);
    CHECK(goofy_server_destroy(s) == GOOFY_OK);
}

void case_deadline()
{
    std::printf("case: deadlines that pass in the queue, before submit, and that hold\n");
    std::fflush(stdout);
    set_fake(30000);
    const std::vector<Tex> tex = make_textures(8, 13);
    goofy_server* s = make_server(64ull << 20, 16ull << 20, 1, 1, 1);
    const uint64_t kernels0 = fake_hip_kernels_run();
    const double now = goofy_server_now_us(s);
    goofy_request r[4] = {req(tex[0], 0), req(tex[1], 1, 0, now + 2000.0), req(tex[2], 2, 0, now + 30e6),
                          req(tex[3], 3, 0, now - 1.0)};
    uint64_t t[4];
    int accepted = 0;
    CHECK(goofy_server_submit_many(s, r, 4, t, &accepted) == GOOFY_OK && accepted == 4);
    const auto got = collect(s, tex, 4);
    CHECK(got.at(t[0]).status == GOOFY_OK);
    CHECK(got.at(t[1]).status == GOOFY_ERR_DEADLINE && got.at(t[1]).dst == nullptr);
    CHECK(got.at(t[1]).t_done_us >= now + 2000.0 && got.at(t[1]).t_done_us < now + 30000.0);  // before request 0 ends
    CHECK(got.at(t[2]).status == GOOFY_OK);
    CHECK(got.at(t[3]).status == GOOFY_ERR_DEADLINE);
    CHECK(goofy_server_wait_idle(s, -1) == GOOFY_OK);
    CHECK(fake_hip_kernels_run() - kernels0 == 2);  // the expired requests never ran
    goofy_server_stats st{};
    CHECK(goofy_server_stats_get(s, &st) == GOOFY_OK);
    CHECK(st.expired == 2 && st.admitted == 2 && st.inflight_bytes == 0);
    CHECK(goofy_server_destroy(s) == GOOFY_OK);
}

void case_wait_idle_needs_release()
{
    std::printf("case: wait_idle with the arena held by unreleased outputs\n");
    std::fflush(stdout);
    set_fake(100);
    std::vector<Tex> tex = make_textures(1, 17);
    const Tex& t = tex[0];
    const size_t out = (t.out_bytes() + 255) / 256 * 256;
    goofy_server* s = make_server(64ull << 20, 2 * out, 2, 1, 1);  // the arena holds two outputs
    uint64_t tk[4];
    for (int i = 0; i < 4; ++i) {
        goofy_request r = req(t, 0);
        CHECK(goofy_server_submit(s, &r, &tk[i]) == GOOFY_OK);
    }
    // two run and complete; their outputs fill the arena; the other two wait
    auto f = std::async(std::launch::async, [s] { return goofy_server_wait_idle(s, -1); });
    CHECK(f.wait_for(std::chrono::seconds(10)) == std::future_status::ready);
    CHECK(f.get() == GOOFY_ERR_BUSY);
    CHECK(goofy_server_wait_idle(s, 50) == GOOFY_ERR_BUSY);
    const auto first = collect(s, tex, 2, true);  // releases let the other two in
    CHECK(goofy_server_wait_idle(s, 10000) == GOOFY_OK);
    const auto second = collect(s, tex, 2, true);
    CHECK(first.size() == 2 && second.size() == 2);
    goofy_server_stats st{};
    CHECK(goofy_server_stats_get(s, &st) == GOOFY_OK);
    CHECK(st.completed == 4 && st.released == 4 && st.inflight_bytes == 0);
    CHECK(goofy_server_destroy(s) == GOOFY_OK);
}

void case_launch_failure_and_fault()
{
    std::printf("case: a launch call that fails, then a kernel that faults its stream\n");
    std::fflush(stdout);
    const std::vector<Tex> tex = make_textures(8, 19);
    for (int fault : {0, 1}) {
        set_fake(200, -1, fault ? -1 : 2, fault ? 2 : -1);
        goofy_server* s = make_server(64ull << 20, 16ull << 20, 1, 1, 1);
        uint64_t tk[6];
        for (int i = 0; i < 6; ++i) {
            goofy_request r = req(tex[i], (uint64_t)i);
            CHECK(goofy_server_submit(s, &r, &tk[i]) == GOOFY_OK);
        }
        const auto got = collect(s, tex, 6, true);  // error completions are released too
        CHECK(got.at(tk[0]).status == GOOFY_OK && got.at(tk[1]).status == GOOFY_OK);
        CHECK(got.at(tk[2]).status == GOOFY_ERR_HIP && got.at(tk[2]).dst != nullptr);
        if (!fault)  // a failed launch call spoils only its own group
            for (int i = 3; i < 6; ++i) CHECK(got.at(tk[i]).status == GOOFY_OK);
        else  // a fault is sticky: the stream's later groups fail too
            for (int i = 3; i < 6; ++i) CHECK(got.at(tk[i]).status == GOOFY_ERR_HIP);
        CHECK(goofy_server_wait_idle(s, -1) == GOOFY_OK);
        goofy_server_stats st{};
        CHECK(goofy_server_stats_get(s, &st) == GOOFY_OK);
        CHECK(st.completed == 6 && st.released == 6 && st.inflight_bytes == 0);
        CHECK(goofy_server_destroy(s) == GOOFY_OK);
    }
    set_fake(0);
}

void case_allocation_failure()
{
    std::printf("case: allocation failures while creating a server\n");
    std::fflush(stdout);
    for (int k : {0, 1, 2}) {  // the arena, then the first slots' device work-lists
        set_fake(0, k);
        goofy_server_config c{};
        c.hbm_budget_bytes = 1 << 20;
        c.streams = 2;
        goofy_server* s = reinterpret_cast<goofy_server*>(0x1);
        CHECK(goofy_server_create(&c, &s) == GOOFY_ERR_OUT_OF_MEMORY && s == nullptr);
    }
    set_fake(0);
}

void case_destroy_in_flight()
{
    std::printf("case: destroy with requests pending and running\n");
    std::fflush(stdout);
    set_fake(2000);
    const std::vector<Tex> tex = make_textures(16, 23);
    goofy_server* s = make_server(64ull << 20, 16ull << 20, 2, 2, 1);
    for (int i = 0; i < 50; ++i) {
        goofy_request r = req(tex[i % tex.size()], (uint64_t)i);
        CHECK(goofy_server_submit(s, &r, nullptr) == GOOFY_OK);
    }
    std::this_thread::sleep_for(std::chrono::milliseconds(3));
    CHECK(goofy_server_destroy(s) == GOOFY_OK);
    set_fake(0);
}

void case_simulator()
{
    std::printf("case: Poisson load generator, every result verified\n");
    std::fflush(stdout);
    set_fake(50);
    const std::vector<Tex> tex = make_textures(24, 29, 64);
    std::vector<goofy_texture_desc> pool;
    std::vector<const unsigned char*> expected;
    for (const Tex& t : tex) {
        pool.push_back(goofy_texture_desc{t.px, nullptr, t.w, t.h, t.pitch, t.fmt});
        expected.push_back(t.expected.data());
    }
    goofy_serve_sim_config c{};
    c.rate_per_s = 4000;
    c.duration_s = 0.3;
    c.seed = 5;
    c.pool = pool.data();
    c.pool_count = (int)pool.size();
    c.pool_expected = expected.data();
    c.verify_every = 1;
    c.server.hbm_budget_bytes = 64ull << 20;
    c.server.streams = 4;
    goofy_serve_sim_result r{};
    CHECK(goofy_serve_simulate(&c, &r) == GOOFY_OK);
    std::printf("  submitted %llu completed %llu p50 %.1f p99 %.1f us, generator lag p99 %.1f us\n",
                (unsigned long long)r.submitted, (unsigned long long)r.completed, r.lat_p50_us, r.lat_p99_us,
                r.gen_lag_p99_us);
    CHECK(r.submitted > 500 && r.completed == r.submitted && r.verified == r.completed && r.mismatches == 0);
    CHECK(r.gen_lag_p50_us >= 0 && r.gen_lag_p99_us >= r.gen_lag_p50_us);
    CHECK(r.lat_p99_us >= r.submit_lat_p99_us - 1e-6 || r.lat_p99_us >= r.submit_lat_p50_us);
    CHECK(r.lat_p50_us > 0 && r.lat_p50_us <= r.lat_p99_us);
    set_fake(0);
}

}  // namespace

int main()
{
    case_concurrent_clients(1);
    case_concurrent_clients(0);
    case_budget_and_queue_limit();
    case_cancel();
    case_deadline();
    case_wait_idle_needs_release();
    case_launch_failure_and_fault();
    case_allocation_failure();
    case_destroy_in_flight();
    case_simulator();
    std::printf("serve fake ok\n");
    return 0;
}
