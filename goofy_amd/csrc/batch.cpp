// batch.cpp -- many textures in one launch.
//
// A plan turns n device-resident textures of uneven sizes into one flattened
// work-list: texture t owns wave tiles [tile_begin(t), tile_begin(t+1)), a
// prefix sum over ceil(blocks(t) / tile).  Tiles never straddle textures, so
// a wave's loads stay contiguous and its format is uniform.  One kernel launch
// walks the whole list (kernels.hip, k_encode_batch), one wave per tile or a
// persistent grid of waves that each take a contiguous run of tiles; the
// launch is captured once into a hipGraph and replayed.  The launch shape
// (blocks per lane, workgroup size, persistent waves per CU, streaming hints)
// comes from an
// autotuner that times the candidates on the plan itself and caches the
// winner per shape class for the process (and, with GOOFY_TUNE_CACHE=<file>,
// across processes).
//
// The reference has no batching: its driver encodes one image with one
// nested loop (GoofyTC/goofy_tc.h:1514-1524).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "goofy_hip.h"
#include "internal.h"

using namespace goofy_impl;

struct goofy_batch {
    int device = -1;
    int num_cus = 0;
    std::string arch;
    std::vector<goofy_texture_desc> textures;  // non-empty textures, caller order
    std::vector<TexDesc> descs;
    TexDesc* d_descs = nullptr;
    size_t d_descs_cap = 0;
    uint32_t ntiles = 0;
    uint64_t blocks = 0, pixels = 0;
    bool alpha = false;  // holds DXT5 / ETC2 RGBA8 textures
    LaunchCfg cfg{};
    hipStream_t capture_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // Recorded after every graph launch, on the caller's stream: the work-list
    // is rewritten or freed only once the launches that read it have finished.
    hipEvent_t launched = nullptr;
    bool launched_pending = false;
};

namespace {

// Waits for the plan's last graph launch (which may be on a caller's
// non-blocking stream, which a synchronous copy does not wait for).
int wait_launched(goofy_batch* b)
{
    if (!b->launched_pending) return GOOFY_OK;
    const hipError_t e = hipEventSynchronize(b->launched);
    if (e != hipSuccess) return fail_hip(e, "hipEventSynchronize(batch launch)");
    b->launched_pending = false;
    return GOOFY_OK;
}

void drop_graph(goofy_batch* b)
{
    if (b->exec) {
        (void)wait_launched(b);
        (void)hipGraphExecDestroy(b->exec);
    }
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Rebuild the tile prefix sums for the plan's tile size and upload them.
int build_worklist(goofy_batch* b)
{
    int rc = wait_launched(b);
    if (rc != GOOFY_OK) return rc;
    const uint64_t tile = (uint64_t)tile_blocks(b->cfg);
    b->descs.clear();
    uint64_t tiles = 0;
    for (const goofy_texture_desc& t : b->textures) {
        // one entry per horizontal stripe of at most max_stripe_brows block
        // rows (a single stripe unless the texture is beyond the kernel's
        // index range)
        const uint32_t bw = t.width / 4, bh = t.height / 4, max_rows = max_stripe_brows(bw);
        for (uint32_t y0 = 0; y0 < bh;) {
            const uint32_t rows = std::min(max_rows, bh - y0);
            TexDesc d{};
            d.src = (uint64_t)(uintptr_t)t.src + (uint64_t)y0 * 4 * t.pitch;
            d.dst = (uint64_t)(uintptr_t)t.dst + (uint64_t)y0 * bw * block_bytes(t.format);
            d.pitch = t.pitch;
            d.bw = bw;
            d.nblocks = rows * bw;
            d.magic = div_magic(bw);
            d.fmt = (uint32_t)t.format;
            d.tile_begin = (uint32_t)tiles;
            tiles += (d.nblocks + tile - 1) / tile;
            if (tiles > 0xFFFFFFFFull / tile) return fail(GOOFY_ERR_TOO_LARGE, "batch has too many tiles");
            b->descs.push_back(d);
            y0 += rows;
        }
    }
    b->ntiles = (uint32_t)tiles;
    const size_t ntex = b->descs.size();
    std::vector<uint8_t> host(worklist_bytes(ntex));
    if (ntex) std::memcpy(host.data(), b->descs.data(), ntex * sizeof(TexDesc));
    uint32_t* tb = worklist_tile_begin(reinterpret_cast<TexDesc*>(host.data()), ntex);
    for (size_t i = 0; i < ntex; ++i) tb[i] = b->descs[i].tile_begin;
    tb[ntex] = b->ntiles;
    const size_t bytes = host.size();
    hipError_t e;
    if (b->d_descs_cap < bytes) {
        if (b->d_descs) (void)hipFree(b->d_descs);
        b->d_descs = nullptr;
        b->d_descs_cap = 0;
        e = hipMalloc((void**)&b->d_descs, bytes);
        if (e != hipSuccess) return fail_hip(e, "hipMalloc(descriptors)");
        b->d_descs_cap = bytes;
    }
    e = hipMemcpy(b->d_descs, host.data(), bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail_hip(e, "hipMemcpy(work-list)");
    drop_graph(b);
    return GOOFY_OK;
}

int capture(goofy_batch* b)
{
    if (b->exec) return GOOFY_OK;
    hipError_t e;
    if (!b->capture_stream) {
        e = hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking);
        if (e != hipSuccess) return fail_hip(e, "hipStreamCreate");
    }
    e = hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return fail_hip(e, "hipStreamBeginCapture");
    hipError_t le = launch_batch(b->cfg, b->d_descs, (uint32_t)b->descs.size(), b->ntiles, b->num_cus,
                                 b->capture_stream, b->alpha);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(b->capture_stream, &g);
    if (le != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail_hip(le, "batch kernel launch (capture)");
    }
    if (e != hipSuccess) return fail_hip(e, "hipStreamEndCapture");
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail_hip(e, "hipGraphInstantiate");
    }
    b->graph = g;
    b->exec = x;
    return GOOFY_OK;
}

// ---- tuning cache ---------------------------------------------------------------

using Key = std::tuple<std::string, int, int, int>;  // arch, log2(blocks), log2(blocks/texture), %DXT1 / 10

std::mutex g_tune_mu;
std::map<Key, LaunchCfg>& tune_cache()
{
    static auto* m = new std::map<Key, LaunchCfg>();
    return *m;
}

Key shape_key(const goofy_batch* b)
{
    uint64_t dxt = 0, alpha = 0;
    for (const auto& t : b->textures) {
        dxt += t.format == GOOFY_FORMAT_DXT1;
        alpha += t.format >= GOOFY_FORMAT_DXT5;
    }
    const int n = (int)std::max<size_t>(1, b->textures.size());
    const int lb = (int)std::log2((double)std::max<uint64_t>(1, b->blocks));
    const int lpt = (int)std::log2((double)std::max<uint64_t>(1, b->blocks / (uint64_t)n));
    // share mix: tenths of DXT1 textures, plus 11 x tenths of alpha textures
    return Key{b->arch, lb, lpt, (int)(dxt * 10 / (uint64_t)n + 11 * (alpha * 10 / (uint64_t)n))};
}

void load_file_cache()
{
    static bool loaded = false;
    if (loaded) return;
    loaded = true;
    const char* path = std::getenv("GOOFY_TUNE_CACHE");
    if (!path) return;
    FILE* f = std::fopen(path, "r");
    if (!f) return;
    char arch[64];
    int lb, lpt, mix, bpl, wg, wpc, nt, order;
    while (std::fscanf(f, "%63s %d %d %d %d %d %d %d %d", arch, &lb, &lpt, &mix, &bpl, &wg, &wpc, &nt, &order) == 9) {
        LaunchCfg c{bpl, wg, wpc, nt, order};
        if (cfg_valid(c)) tune_cache()[Key{arch, lb, lpt, mix}] = c;
    }
    std::fclose(f);
}

void save_file_cache()
{
    const char* path = std::getenv("GOOFY_TUNE_CACHE");
    if (!path) return;
    FILE* f = std::fopen(path, "w");
    if (!f) return;
    for (const auto& kv : tune_cache()) {
        const LaunchCfg& c = kv.second;
        std::fprintf(f, "%s %d %d %d %d %d %d %d %d\n", std::get<0>(kv.first).c_str(), std::get<1>(kv.first),
                     std::get<2>(kv.first), std::get<3>(kv.first), c.blocks_per_lane, c.wg_waves, c.waves_per_cu,
                     c.nontemporal, c.tile_order);
    }
    std::fclose(f);
}

int check_texture(const goofy_texture_desc& t, size_t i)
{
    char where[64];
    std::snprintf(where, sizeof where, "texture %zu: ", i);
    if (!format_valid(t.format)) return fail(GOOFY_ERR_INVALID, std::string(where) + "unknown format");
    if (t.width % 4 != 0) return fail(GOOFY_ERR_WIDTH, std::string(where) + "width must be a multiple of 4");
    if (t.height % 4 != 0) return fail(GOOFY_ERR_HEIGHT, std::string(where) + "height must be a multiple of 4");
    if (t.width == 0 || t.height == 0) return GOOFY_OK;
    if (!t.src || !t.dst) return fail(GOOFY_ERR_INVALID, std::string(where) + "null buffer");
    if (t.pitch < (size_t)t.width * 4 || t.pitch % 16 != 0)
        return fail(GOOFY_ERR_INVALID, std::string(where) + "pitch must be >= 4*width and a multiple of 16");
    if ((uintptr_t)t.src % 16 != 0 || (uintptr_t)t.dst % block_bytes(t.format) != 0)
        return fail(GOOFY_ERR_INVALID, std::string(where) + "src must be 16-byte and dst block-size aligned");
    if (max_stripe_brows(t.width / 4) == 0)  // larger textures become several work-list entries
        return fail(GOOFY_ERR_TOO_LARGE, std::string(where) + "too wide (4,194,304 pixels or more per row)");
    return GOOFY_OK;
}

}  // namespace

extern "C" {

int goofy_batch_create(const goofy_texture_desc* textures, size_t count, const goofy_launch_config* ucfg,
                       goofy_batch** out)
{
    clear_error();
    if (!out || (count && !textures)) return fail(GOOFY_ERR_INVALID, "null argument");
    *out = nullptr;
    for (size_t i = 0; i < count; ++i) {
        const int rc = check_texture(textures[i], i);
        if (rc != GOOFY_OK) return rc;
    }
    DeviceInfo di;
    int rc = device_info(&di);
    if (rc != GOOFY_OK) return rc;
    goofy_batch* b = new goofy_batch();
    b->device = di.device;
    b->num_cus = di.num_cus;
    b->arch = di.arch;
    for (size_t i = 0; i < count; ++i) {
        const goofy_texture_desc& t = textures[i];
        if (t.width == 0 || t.height == 0) continue;
        b->textures.push_back(t);
        b->alpha |= t.format >= GOOFY_FORMAT_DXT5;
        b->blocks += (uint64_t)(t.width / 4) * (t.height / 4);
        b->pixels += (uint64_t)t.width * t.height;
    }
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        load_file_cache();
        auto it = tune_cache().find(shape_key(b));
        b->cfg = it != tune_cache().end() ? it->second : auto_cfg(b->blocks, b->num_cus);
    }
    if (ucfg) {
        LaunchCfg c = to_cfg(*ucfg);
        if (!cfg_valid(c)) {
            delete b;
            return fail(GOOFY_ERR_INVALID, "invalid launch config");
        }
        b->cfg = c;
    }
    rc = build_worklist(b);
    if (rc != GOOFY_OK) {
        goofy_batch_destroy(b);
        return rc;
    }
    *out = b;
    return GOOFY_OK;
}

int goofy_batch_launch(goofy_batch* b, void* stream)
{
    clear_error();
    if (!b) return fail(GOOFY_ERR_INVALID, "null batch");
    if (b->ntiles == 0) return GOOFY_OK;
    int rc = capture(b);
    if (rc != GOOFY_OK) return rc;
    hipError_t e = hipGraphLaunch(b->exec, (hipStream_t)stream);
    if (e != hipSuccess) return fail_hip(e, "hipGraphLaunch");
    if (!b->launched && (e = hipEventCreateWithFlags(&b->launched, hipEventDisableTiming)) != hipSuccess)
        return fail_hip(e, "hipEventCreate");
    if ((e = hipEventRecord(b->launched, (hipStream_t)stream)) != hipSuccess) return fail_hip(e, "hipEventRecord");
    b->launched_pending = true;
    return GOOFY_OK;
}

int goofy_batch_launch_per_texture(goofy_batch* b, void* stream)
{
    clear_error();
    if (!b) return fail(GOOFY_ERR_INVALID, "null batch");
    for (const goofy_texture_desc& t : b->textures) {
        // horizontal stripes of at most max_stripe_brows block rows, as in
        // the work-list: beyond that a launch's block index or its division
        // magic would be out of range
        const uint32_t bw = t.width / 4, bh = t.height / 4, max_rows = max_stripe_brows(bw);
        for (uint32_t y0 = 0; y0 < bh;) {
            const uint32_t rows = std::min(max_rows, bh - y0);
            const uint32_t nb = rows * bw;
            const LaunchCfg c = auto_cfg(nb, b->num_cus);
            hipError_t e = launch_plain(t.format, c, static_cast<const uint8_t*>(t.src) + (size_t)y0 * 4 * t.pitch,
                                        t.pitch, static_cast<uint8_t*>(t.dst) + (size_t)y0 * bw * block_bytes(t.format),
                                        bw, nb, (hipStream_t)stream);
            if (e != hipSuccess) return fail_hip(e, "per-texture launch");
            y0 += rows;
        }
    }
    return GOOFY_OK;
}

int goofy_batch_get_config(const goofy_batch* b, goofy_launch_config* cfg)
{
    clear_error();
    if (!b || !cfg) return fail(GOOFY_ERR_INVALID, "null argument");
    *cfg = to_public(b->cfg);
    return GOOFY_OK;
}

int goofy_batch_set_config(goofy_batch* b, const goofy_launch_config* cfg)
{
    clear_error();
    if (!b || !cfg) return fail(GOOFY_ERR_INVALID, "null argument");
    LaunchCfg c = to_cfg(*cfg);
    if (!cfg_valid(c)) return fail(GOOFY_ERR_INVALID, "invalid launch config");
    const bool retile = c.blocks_per_lane != b->cfg.blocks_per_lane;
    b->cfg = c;
    if (retile) return build_worklist(b);
    drop_graph(b);
    return GOOFY_OK;
}

int goofy_batch_autotune(goofy_batch* b, void* stream, int reps, goofy_launch_config* chosen, float* best_ms)
{
    clear_error();
    if (!b) return fail(GOOFY_ERR_INVALID, "null batch");
    if (reps < 1) reps = 1;
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0, e1;
    hipError_t e = hipEventCreate(&e0);
    if (e != hipSuccess) return fail_hip(e, "hipEventCreate");
    e = hipEventCreate(&e1);
    if (e != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail_hip(e, "hipEventCreate");
    }
    const LaunchCfg start = b->cfg;
    LaunchCfg best = b->cfg;
    float best_t = 1e30f;
    int rc = GOOFY_OK;
    const int bpls[] = {1, 2, 4}, wgs[] = {2, 4, 8}, wpcs[] = {0, 8, 16, 32}, nts[] = {0, 1}, orders[] = {0, 1};
    for (int bpl : bpls) {
        b->cfg.blocks_per_lane = bpl;
        if ((rc = build_worklist(b)) != GOOFY_OK) break;
        for (int wg : wgs)
            for (int wpc : wpcs)
                for (int nt : nts)
                    for (int order : orders) {
                        b->cfg = LaunchCfg{bpl, wg, wpc, nt, order};
                        // one warm-up launch, then the median of `reps` timed ones
                        e = launch_batch(b->cfg, b->d_descs, (uint32_t)b->descs.size(), b->ntiles, b->num_cus, s, b->alpha);
                        if (e != hipSuccess) { rc = fail_hip(e, "autotune launch"); goto done; }
                        std::vector<float> ts;
                        for (int r = 0; r < reps; ++r) {
                            (void)hipEventRecord(e0, s);
                            e = launch_batch(b->cfg, b->d_descs, (uint32_t)b->descs.size(), b->ntiles, b->num_cus, s, b->alpha);
                            (void)hipEventRecord(e1, s);
                            if (e != hipSuccess) { rc = fail_hip(e, "autotune launch"); goto done; }
                            e = hipEventSynchronize(e1);
                            if (e != hipSuccess) { rc = fail_hip(e, "hipEventSynchronize"); goto done; }
                            float ms = 0;
                            (void)hipEventElapsedTime(&ms, e0, e1);
                            ts.push_back(ms);
                        }
                        std::sort(ts.begin(), ts.end());
                        const float med = ts[ts.size() / 2];
                        if (med < best_t) {
                            best_t = med;
                            best = b->cfg;
                        }
                    }
    }
done:
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != GOOFY_OK) {
        b->cfg = start;
        (void)build_worklist(b);
        return rc;
    }
    b->cfg = best;
    if ((rc = build_worklist(b)) != GOOFY_OK) return rc;
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        tune_cache()[shape_key(b)] = best;
        save_file_cache();
    }
    if (chosen) *chosen = to_public(best);
    if (best_ms) *best_ms = best_t;
    return GOOFY_OK;
}

int goofy_batch_info(const goofy_batch* b, uint64_t* textures, uint64_t* blocks, uint64_t* tiles, uint64_t* pixels)
{
    clear_error();
    if (!b) return fail(GOOFY_ERR_INVALID, "null batch");
    if (textures) *textures = b->textures.size();
    if (blocks) *blocks = b->blocks;
    if (tiles) *tiles = b->ntiles;
    if (pixels) *pixels = b->pixels;
    return GOOFY_OK;
}

int goofy_batch_destroy(goofy_batch* b)
{
    clear_error();
    if (!b) return GOOFY_OK;
    (void)wait_launched(b);
    if (b->launched) (void)hipEventDestroy(b->launched);
    drop_graph(b);
    if (b->capture_stream) (void)hipStreamDestroy(b->capture_stream);
    if (b->d_descs) (void)hipFree(b->d_descs);
    delete b;
    return GOOFY_OK;
}

}  // extern "C"
