/*
 * goofy_hip.h -- C-ABI of libgoofy_hip.so, the MI355X (gfx950) Goofy block
 * encoder.  Plain pointers and sizes only; every function returns GOOFY_OK (0)
 * or a negative goofy_status code, and goofy_hip_last_error() gives the
 * message of the calling thread's last failure.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   goofy::compressDXT1 / goofy::compressETC1   GoofyTC/goofy_tc.h:10-13, :1497-1557
 *   the per-4-block kernel goofySimdEncode<C>   GoofyTC/goofy_tc.h:1069-1494
 *   the harness plugin type CompressFunc_t      Src/main.cpp:644
 */
#ifndef GOOFY_HIP_H
#define GOOFY_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GOOFY_HIP_API __attribute__((visibility("default")))

/* DXT1 and ETC1S blocks are 8 bytes.  DXT5 (BC3) and ETC2 RGBA8 (EAC alpha)
 * blocks are 16: an alpha block, then the DXT1 / ETC1S block of the same
 * pixels, bit-identical to the 8-byte formats.  The reference has no alpha
 * encoder (its README names one as the next step, "alpha directly instead of
 * brightness"); the alpha blocks here put every pixel on its nearest palette
 * level (BC3: a0 = max, a1 = min alpha; EAC: table 9, multiplier
 * max(1, ceil((max-min)/19)), base on the minimum or, near 255, the maximum).
 * The host drop-in, the device entry points, the decoders, quality sums and
 * containers, batches and the server take all four. */
typedef enum goofy_format {
    GOOFY_FORMAT_DXT1 = 0,
    GOOFY_FORMAT_ETC1 = 1,
    GOOFY_FORMAT_DXT5 = 2,
    GOOFY_FORMAT_ETC2 = 3
} goofy_format;

/* Bytes per 4x4 block of a format (8 or 16; 0 for an unknown format). */
GOOFY_HIP_API unsigned int goofy_format_block_bytes(int format);

typedef enum goofy_status {
    GOOFY_OK = 0,
    GOOFY_ERR_WIDTH = -1,          /* width % 16 != 0 (drop-in) or width % 4 != 0 (device/batch API) */
    GOOFY_ERR_HEIGHT = -2,         /* height % 4 != 0 */
    GOOFY_ERR_HIP = -3,            /* HIP runtime error (message has the HIP error string) */
    GOOFY_ERR_INVALID = -4,        /* null pointer, bad format, stride/pitch too small or misaligned, bad config */
    GOOFY_ERR_OUT_OF_MEMORY = -5,  /* device or pinned allocation failed */
    GOOFY_ERR_NO_DEVICE = -6,      /* no GPU visible */
    GOOFY_ERR_TOO_LARGE = -7,      /* rows of 4,194,304 pixels or more (larger images of narrower rows run
                                      as horizontal stripes); probe and serving: blocks * blocks-per-row
                                      >= 2^40 */
    GOOFY_ERR_BUSY = -8,           /* serving: admission refused (queue full), request already launched
                                      (cancel), or pending requests wait for outputs to be released (wait_idle) */
    GOOFY_ERR_CANCELLED = -9,      /* serving: completion status of a request cancelled before its launch */
    GOOFY_ERR_DEADLINE = -10       /* serving: completion status of a request whose deadline passed before its
                                      launch (it was never launched) */
} goofy_status;

/* Message for the calling thread's most recent failure ("" if none). */
GOOFY_HIP_API const char* goofy_hip_last_error(void);
GOOFY_HIP_API const char* goofy_hip_status_string(int status);
/* Library version string and the gfx target it was built for. */
GOOFY_HIP_API const char* goofy_hip_version(void);
/* Number of visible GPUs (0 when none; never fails). */
GOOFY_HIP_API int goofy_hip_device_count(void);

/* ---- drop-in, host memory ----------------------------------------------
 * Same contract as goofy::compressDXT1/compressETC1 (GoofyTC/goofy_tc.h:1497-1557):
 * host `input` RGBA8 with `stride` bytes per row, host `result` receives
 * (width/4)*(height/4) tightly packed blocks (8 bytes each; 16 for DXT5 and
 * ETC2 RGBA8).  -1 if width % 16, then -2 if
 * height % 4; width or height 0 writes nothing and returns 0.  Synchronous:
 * stages through device memory on a per-thread stream and returns when
 * `result` holds the blocks.  Thread-safe. */
GOOFY_HIP_API int goofy_hip_compress(int format, unsigned char* result, const unsigned char* input,
                                     unsigned int width, unsigned int height, unsigned int stride);
/* The same entry points with the exact CompressFunc_t signature
 * (Src/main.cpp:644), so a harness can swap them for the reference's.  The
 * DXT5 / ETC2 ones have no reference encoder to replace; the harness's
 * decompressDXT5 / decompressETC2 (Src/main.cpp:579-631) read their output. */
GOOFY_HIP_API int goofy_hip_compress_dxt1(unsigned char* result, const unsigned char* input, unsigned int width,
                                          unsigned int height, unsigned int stride);
GOOFY_HIP_API int goofy_hip_compress_etc1(unsigned char* result, const unsigned char* input, unsigned int width,
                                          unsigned int height, unsigned int stride);
GOOFY_HIP_API int goofy_hip_compress_dxt5(unsigned char* result, const unsigned char* input, unsigned int width,
                                          unsigned int height, unsigned int stride);
GOOFY_HIP_API int goofy_hip_compress_etc2(unsigned char* result, const unsigned char* input, unsigned int width,
                                          unsigned int height, unsigned int stride);

/* Both formats of a host image from one pass: DXT1 blocks into result_dxt1
 * and ETC1S blocks into result_etc1 (each (width/4)*(height/4)*8 bytes),
 * bit-identical to goofy_hip_compress_dxt1 and _etc1.  The image crosses PCIe
 * once instead of twice.  Same checks, codes and synchronous behaviour as the
 * drop-in.  No counterpart in the reference (one format per call there). */
GOOFY_HIP_API int goofy_hip_compress_dual(unsigned char* result_dxt1, unsigned char* result_etc1,
                                          const unsigned char* input, unsigned int width, unsigned int height,
                                          unsigned int stride);
/* The same for the 16-byte formats: DXT5 into result_dxt5 and ETC2 RGBA8 into
 * result_etc2 (each (width/4)*(height/4)*16 bytes) from one transfer and one
 * pass, bit-identical to goofy_hip_compress_dxt5 and _etc2. */
GOOFY_HIP_API int goofy_hip_compress_dual_alpha(unsigned char* result_dxt5, unsigned char* result_etc2,
                                                const unsigned char* input, unsigned int width, unsigned int height,
                                                unsigned int stride);

/* ---- device memory, asynchronous ------------------------------------------
 * Encodes a device-resident RGBA8 image into device-resident blocks on
 * `stream` (a hipStream_t, NULL = default stream) and returns after the
 * launch.  Per-block semantics: width and height multiples of 4 (width % 16
 * is NOT required), `pitch` >= 4*width and a multiple of 16, `d_input`
 * 16-byte aligned, `d_result` aligned to the block size (8 or 16 bytes).
 * Output is bit-identical to the reference on every block (for DXT5 / ETC2
 * RGBA8: the colour half; the alpha half follows the rule above). */
GOOFY_HIP_API int goofy_hip_compress_device(int format, void* d_result, const void* d_input, unsigned int width,
                                            unsigned int height, size_t pitch, void* stream);

/* Launch shape used by the device and batch entry points. */
typedef struct goofy_launch_config {
    int blocks_per_lane; /* 1, 2 or 4 */
    int wg_waves;        /* waves per workgroup: 1, 2, 4 or 8 */
    int waves_per_cu;    /* persistent grid: waves per CU (0 = one wave per tile) */
    int nontemporal;     /* 1 = streaming loads/stores */
    int tile_order;      /* 0 = linear wave->tile order; 1 = XCD-contiguous: the
                            workgroups the dispatcher places on each of the 8
                            XCDs (workgroup b on XCD b % 8) take one contiguous
                            eighth of the tiles */
} goofy_launch_config;

/* As goofy_hip_compress_device with an explicit launch shape.  NULL = the
 * default shape, sized to the launch: one block per lane, streaming loads and
 * stores, linear tile order, and the largest workgroup of 8, 4, 2 or 1 waves
 * that still gives at least two workgroups per CU. */
GOOFY_HIP_API int goofy_hip_compress_device_cfg(int format, void* d_result, const void* d_input, unsigned int width,
                                                unsigned int height, size_t pitch, void* stream,
                                                const goofy_launch_config* cfg);

/* As goofy_hip_compress_device_cfg, with two hipEvent_t handles that the
 * kernel's dispatch packet records at its start and end (either may be NULL):
 * hipEventElapsedTime(start, stop) is then the kernel's own execution time. */
GOOFY_HIP_API int goofy_hip_compress_device_timed(int format, void* d_result, const void* d_input,
                                                  unsigned int width, unsigned int height, size_t pitch, void* stream,
                                                  const goofy_launch_config* cfg, void* start_event, void* stop_event);

/* One device-resident image into both formats in one pass: each block is
 * read and classified once and packed as DXT1 into d_dxt1 and as ETC1S into
 * d_etc1 (each (width/4)*(height/4)*8 bytes, 8-byte aligned), bit-identical to
 * the two single-format encodes.  Reads the image once: 80 bytes of traffic
 * per block instead of 2 x 72 for two launches.  No counterpart in the
 * reference, whose API encodes one format per call; for pipelines that ship
 * both.  Same constraints as goofy_hip_compress_device; events (either may be
 * NULL) as in goofy_hip_compress_device_timed. */
GOOFY_HIP_API int goofy_hip_compress_device_dual(void* d_dxt1, void* d_etc1, const void* d_input, unsigned int width,
                                                 unsigned int height, size_t pitch, void* stream, void* start_event,
                                                 void* stop_event);
/* The same for the 16-byte formats: DXT5 into d_dxt5 and ETC2 RGBA8 into
 * d_etc2 (each (width/4)*(height/4)*16 bytes, 16-byte aligned); the colour
 * classification and the alpha rows of a block are computed once.  96 bytes
 * of traffic per block instead of 2 x 80. */
GOOFY_HIP_API int goofy_hip_compress_device_dual_alpha(void* d_dxt5, void* d_etc2, const void* d_input,
                                                       unsigned int width, unsigned int height, size_t pitch,
                                                       void* stream, void* start_event, void* stop_event);

/* Launch-shape autotuner for one image shape.  Times the encode kernel under
 * each candidate shape (1, 2 or 4 blocks per lane; workgroups of 1, 2, 4 or 8
 * waves; both tile orders; streaming loads and stores) on the caller's
 * buffers -- d_result ends up holding the image's encoding -- and returns the
 * shape with the lowest median kernel time over `reps` timed launches, and
 * that time.  Pass the shape to goofy_hip_compress_device_cfg for later
 * images of the same size and pitch.  Synchronises `stream`. */
GOOFY_HIP_API int goofy_hip_tune_device(int format, void* d_result, const void* d_input, unsigned int width,
                                        unsigned int height, size_t pitch, void* stream, int reps,
                                        goofy_launch_config* best, float* best_ms);

/* ---- decode and round-trip quality ------------------------------------------
 * Decoders with the semantics of the reference's quality harness
 * (Src/decoder.cpp: BC1 :819-871, ETC1 :388-678, BC3 alpha :874-925 with the
 * colour block in 4-colour mode :945-956, EAC alpha :679-713).  `d_rgba`
 * receives RGBA8 rows of `pitch` bytes (multiple of 16, >= 4*width, 16-byte
 * aligned); `d_blocks` holds (width/4)*(height/4) tightly packed blocks of the
 * format's size, aligned to it.  Both ETC formats decode every ETC2 colour
 * mode, as the harness's decoder does (Src/decoder.cpp:716-773): a
 * differential block whose red, green or blue sum leaves 0..31 is a T, H or
 * planar block (:270-386).  DXT1 and ETC1S decode with alpha 255. */
GOOFY_HIP_API int goofy_hip_decode_device(int format, void* d_rgba, const void* d_blocks, unsigned int width,
                                          unsigned int height, size_t pitch, void* stream);

/* Round-trip error of encoded blocks against their source image, with the
 * definitions of the reference harness's getMsePsnr (Src/main.cpp:403-469):
 * a pixel is read as a little-endian word, "r" = bits 16..23, "g" = 8..15,
 * "b" = 0..7, "a" = 24..31; luma = 0.21 r + 0.72 g + 0.07 b; PSNR =
 * 10 log10(peak^2 / mse) with peak 255, and 768 for the summed-RGB figure;
 * an MSE below 1e-7 gives DBL_MAX.  Synchronous on `stream`.  Uses a
 * 60 KiB device scratch buffer and 40 bytes of pinned host memory, into which
 * the last kernel writes the sums, from a per-device pool (allocated on first
 * use, never freed, one pair per concurrent call), so a call never waits on
 * other streams. */
typedef struct goofy_quality {
    uint64_t pixels;
    uint64_t sse_byte[4]; /* squared error of bytes 0..3 of each pixel, summed */
    double sse_luma;
    double mse_r, mse_g, mse_b, mse_a, mse_max, mse_rgb, mse_y;
    double psnr_r, psnr_g, psnr_b, psnr_a, psnr_min, psnr_rgb, psnr_y;
} goofy_quality;

GOOFY_HIP_API int goofy_hip_quality_device(int format, const void* d_blocks, const void* d_source, unsigned int width,
                                           unsigned int height, size_t pitch, void* stream, goofy_quality* out);

/* goofy_hip_quality_device without allocating.  No counterpart in the
 * reference: for callers that cannot allocate device or pinned memory (tight
 * serving loops, harnesses that forbid hipMalloc).  The kernels' partial and
 * final sums live in a fixed 61,504-byte device array of the library, one
 * copy per GPU held until the process ends; the sums reach `out` through a
 * plain device-to-host copy, and the call synchronises `stream`.  It calls no
 * hipMalloc, hipFree or pinned allocation.  Same checks, limits and results
 * as goofy_hip_quality_device (any image that call takes fits the array:
 * larger images run as stripes whose sums accumulate in it); every call
 * overwrites the array, so calls that overlap on one GPU (from several
 * threads, or on several streams) race and must be serialised by the caller. */
GOOFY_HIP_API int goofy_hip_quality_device_noalloc(int format, const void* d_blocks, const void* d_source,
                                                   unsigned int width, unsigned int height, size_t pitch, void* stream,
                                                   goofy_quality* out);

/* getMsePsnr (Src/main.cpp:403-469) between two device RGBA8 images of the
 * same size, e.g. a source and its quality floor below: same figures and
 * labels as goofy_hip_quality_device.  width % 4 == 0; rows 16-byte aligned.
 * Synchronises `stream`. */
GOOFY_HIP_API int goofy_hip_quality_images_device(const void* d_a, size_t pitch_a, const void* d_b, size_t pitch_b,
                                                  unsigned int width, unsigned int height, void* stream,
                                                  goofy_quality* out);

/* The harness's quality floor, "Baseline (Downsample x2)" in RGB565
 * (generateDownsamlpedRgb565Test, Src/main.cpp:501-555, called at :900):
 * every 2x2 cell of the source becomes one colour, its rounded box average
 * cut to 5/6/6 bits and widened back to 8 bits as the harness does it,
 * alpha 255; the output has the source's size.  Bit-exact with the harness.
 * width % 4 == 0 (the harness needs even sizes; its images have width % 16),
 * height % 2 == 0; pitches >= 4*width and multiples of 16; 16-byte aligned.
 * Returns after the launch.  The _timed form records the two events around
 * the kernel's own dispatch, as goofy_hip_compress_device_timed. */
GOOFY_HIP_API int goofy_hip_rgb565_baseline_device(void* d_dst, size_t dst_pitch, const void* d_src, size_t src_pitch,
                                                   unsigned int width, unsigned int height, void* stream);
GOOFY_HIP_API int goofy_hip_rgb565_baseline_device_timed(void* d_dst, size_t dst_pitch, const void* d_src,
                                                         size_t src_pitch, unsigned int width, unsigned int height,
                                                         void* stream, void* start_event, void* stop_event);

/* Roofline calibration: the plain encode kernel's memory traffic (same
 * addresses, tiles of 64 blocks, one per wave, and streaming hints; 64 bytes
 * read and 8 written per block) with no encode arithmetic, in workgroups of
 * wg_waves waves (1, 2, 4 or 8; 0 = 8).  Its best time over the workgroup
 * sizes is the floor for the encode kernels on the device at hand.  Events as
 * in goofy_hip_compress_device_timed. */
GOOFY_HIP_API int goofy_hip_traffic_probe(void* d_out, const void* d_input, unsigned int width, unsigned int height,
                                          size_t pitch, int wg_waves, void* stream, void* start_event,
                                          void* stop_event);

/* ---- batched encode: many textures, one launch -----------------------------
 * A plan flattens the block work-lists of n device-resident textures (each
 * its own size, pitch and format) into one wave-tile list and encodes all of
 * them with one kernel launch, captured once into a hipGraph.  The launch
 * runs one wave per tile, or a persistent grid (waves_per_cu > 0) in which
 * each wave takes a contiguous run of tiles; the autotuner times both. */
typedef struct goofy_texture_desc {
    const void* src;     /* device RGBA8, 16-byte aligned */
    void* dst;           /* device, (width/4)*(height/4) blocks, aligned to the block size */
    unsigned int width;  /* multiple of 4 */
    unsigned int height; /* multiple of 4 */
    size_t pitch;        /* bytes between rows: >= 4*width, multiple of 16 */
    int format;          /* goofy_format */
} goofy_texture_desc;

typedef struct goofy_batch goofy_batch;

/* cfg NULL = the tuned shape recorded for a batch like this one (see
 * goofy_batch_autotune), else the default shape sized to the whole batch.
 * The plan keeps its own device copy of the descriptors (the textures
 * themselves stay owned by the caller). */
GOOFY_HIP_API int goofy_batch_create(const goofy_texture_desc* textures, size_t count,
                                     const goofy_launch_config* cfg, goofy_batch** out);
/* Enqueue the whole batch on `stream` (replays the captured graph). */
GOOFY_HIP_API int goofy_batch_launch(goofy_batch* batch, void* stream);
/* Same work as individual launches (one plain kernel per texture, or per
 * horizontal stripe for textures beyond one launch's index range), for
 * comparison. */
GOOFY_HIP_API int goofy_batch_launch_per_texture(goofy_batch* batch, void* stream);
/* Time the candidate launch shapes on this batch (HIP events, `reps` runs
 * each), keep the fastest in the plan and the process-wide tuning cache, and
 * report it.  `best_ms` may be NULL. */
GOOFY_HIP_API int goofy_batch_autotune(goofy_batch* batch, void* stream, int reps, goofy_launch_config* chosen,
                                       float* best_ms);
GOOFY_HIP_API int goofy_batch_get_config(const goofy_batch* batch, goofy_launch_config* cfg);
/* set_config, autotune and destroy first wait for the plan's earlier
 * goofy_batch_launch calls to finish (on whatever stream they were issued),
 * since they rewrite or free the device work-list those launches read. */
GOOFY_HIP_API int goofy_batch_set_config(goofy_batch* batch, const goofy_launch_config* cfg);
/* Totals of the plan: textures, 4x4 blocks, wave tiles, pixels. */
GOOFY_HIP_API int goofy_batch_info(const goofy_batch* batch, uint64_t* textures, uint64_t* blocks, uint64_t* tiles,
                                   uint64_t* pixels);
GOOFY_HIP_API int goofy_batch_destroy(goofy_batch* batch);

/* ---- serving: continuous batching under an HBM budget ------------------------
 * A server owns a scheduler thread, `streams` HIP streams and an output arena
 * in HBM.  Clients submit requests that point at device-resident RGBA8
 * textures; the scheduler admits pending requests in arrival order while the
 * in-flight byte budget (input + output bytes of admitted, unreleased
 * requests) and the arena allow, packs every group of newly admitted requests
 * into one batch-kernel launch on a free stream, and keeps launching while
 * earlier groups are still running, so work of different arrivals overlaps
 * on the device.  A request's blocks stay in the arena until released. */
typedef struct goofy_server_config {
    size_t hbm_budget_bytes;    /* in-flight budget (0 = 32 GiB) */
    size_t arena_bytes;         /* output arena size (0 = budget / 4) */
    int streams;                /* concurrent launches in flight (0 = 4) */
    int max_batch_textures;     /* textures per launch (0 = 256) */
    int max_pending;            /* queued, not yet admitted requests (0 = 65536) */
    int batching;               /* 1 = one launch per admitted group, 0 = one plain launch per request */
    goofy_launch_config launch; /* kernel shape (all zero = the default shape, sized per launch) */
} goofy_server_config;

typedef struct goofy_request {
    const void* src;     /* device RGBA8, 16-byte aligned */
    unsigned int width;  /* multiple of 4 */
    unsigned int height; /* multiple of 4 */
    size_t pitch;        /* >= 4*width, multiple of 16 */
    int format;          /* goofy_format */
    uint64_t user_tag;
    double arrival_us;   /* server clock (goofy_server_now_us) when the request arrived at the caller,
                            0 = its submit time; completions report it, and latency is measured from it */
    double deadline_us;  /* server clock; still unlaunched then, the request completes with
                            GOOFY_ERR_DEADLINE and is never launched (0 = no deadline) */
} goofy_request;

/* Completions with status GOOFY_OK or GOOFY_ERR_HIP hold an output range (and
 * its budget share) until goofy_server_release; GOOFY_ERR_CANCELLED and
 * GOOFY_ERR_DEADLINE completions never held one (dst NULL) and take no
 * release. */
typedef struct goofy_completion {
    uint64_t ticket;
    uint64_t user_tag;
    int status;          /* GOOFY_OK or an error code */
    void* dst;           /* device address of the blocks (valid until released) */
    size_t dst_bytes;
    double t_arrival_us; /* server clock, microseconds: the request's arrival_us (or its submit time) */
    double t_submit_us;
    double t_launch_us;  /* 0 when never launched */
    double t_done_us;    /* completion observed by the scheduler */
} goofy_completion;

typedef struct goofy_server_stats {
    uint64_t submitted, admitted, completed, released, rejected;
    uint64_t launches, pending, inflight_requests;
    uint64_t inflight_bytes, peak_inflight_bytes, budget_bytes;
    uint64_t cancelled, expired; /* completed with GOOFY_ERR_CANCELLED / GOOFY_ERR_DEADLINE */
} goofy_server_stats;

typedef struct goofy_server goofy_server;

GOOFY_HIP_API int goofy_server_create(const goofy_server_config* cfg, goofy_server** out);
/* Thread-safe.  GOOFY_ERR_BUSY when the pending queue is full, GOOFY_ERR_INVALID
 * for a malformed request or one larger than the budget or the arena. */
GOOFY_HIP_API int goofy_server_submit(goofy_server* server, const goofy_request* request, uint64_t* ticket);
/* Submits requests in order under one lock; stops at the first one refused
 * (its status is returned) and sets *accepted to the number taken.  tickets[i]
 * receives the ticket of request i when accepted (tickets may be NULL). */
GOOFY_HIP_API int goofy_server_submit_many(goofy_server* server, const goofy_request* requests, int count,
                                           uint64_t* tickets, int* accepted);
/* Up to `max` completions, oldest first; *count receives how many. */
GOOFY_HIP_API int goofy_server_poll(goofy_server* server, goofy_completion* out, int max, int* count);
/* Frees the request's output range and its budget share. */
GOOFY_HIP_API int goofy_server_release(goofy_server* server, uint64_t ticket);
/* Releases many tickets under one lock; returns the first error, if any. */
GOOFY_HIP_API int goofy_server_release_many(goofy_server* server, const uint64_t* tickets, int count);
/* Cancels a request that has not been launched: it completes at once with
 * GOOFY_ERR_CANCELLED.  GOOFY_ERR_BUSY when it was already launched (it runs
 * to completion), GOOFY_ERR_INVALID for an unknown or finished ticket. */
GOOFY_HIP_API int goofy_server_cancel(goofy_server* server, uint64_t ticket);
/* Waits until every submitted request has completed (timeout_ms < 0: no
 * limit).  GOOFY_ERR_BUSY at the timeout, or as soon as nothing runs and the
 * oldest pending request waits for budget or arena space that only releases
 * of completed outputs can give back (poll and release, then wait again). */
GOOFY_HIP_API int goofy_server_wait_idle(goofy_server* server, int timeout_ms);
GOOFY_HIP_API int goofy_server_stats_get(goofy_server* server, goofy_server_stats* stats);
GOOFY_HIP_API double goofy_server_now_us(const goofy_server* server);
GOOFY_HIP_API int goofy_server_destroy(goofy_server* server);

/* Load generator around a server: Poisson arrivals at `rate_per_s` (0 =
 * closed loop: keep `closed_loop_depth` requests outstanding) for
 * `duration_s`, each request one texture drawn uniformly from `pool`
 * (device-resident; dst fields ignored) and stamped with its scheduled
 * arrival time, so that latency includes the generator's own lag.  Completed
 * outputs are released at once; every `verify_every`-th is first copied back
 * and compared with `pool_expected[i]` (host bytes of that texture encoded
 * alone; NULL = no check). */
typedef struct goofy_serve_sim_config {
    double rate_per_s;
    double duration_s;
    uint64_t seed;
    int closed_loop_depth;
    const goofy_texture_desc* pool;
    int pool_count;
    const unsigned char* const* pool_expected;
    int verify_every;
    goofy_server_config server;
} goofy_serve_sim_config;

typedef struct goofy_serve_sim_result {
    uint64_t submitted, completed, rejected, verified, mismatches, launches;
    double wall_s, megapixels, mps, offered_mps;
    double lat_mean_us, lat_p50_us, lat_p90_us, lat_p99_us, lat_max_us; /* scheduled arrival -> done */
    double queue_p50_us, queue_p99_us; /* scheduled arrival -> launch */
    double mean_textures_per_launch;
    uint64_t peak_inflight_bytes;
    double gen_lag_p50_us, gen_lag_p99_us, gen_lag_max_us; /* scheduled arrival -> submit (generator lag) */
    double submit_lat_p50_us, submit_lat_p99_us;          /* submit -> done */
} goofy_serve_sim_result;

GOOFY_HIP_API int goofy_serve_simulate(const goofy_serve_sim_config* cfg, goofy_serve_sim_result* result);

/* ---- containers ----------------------------------------------------------------
 * DDS (DXT1, DXT5) and KTX1 (any of the four formats) files around
 * (width/4)*(height/4) blocks, with the headers the reference's harness writes
 * (saveDds / saveKtx, Src/main.cpp:154-220; format codes :64-73).  Host memory
 * only; no GPU needed.  goofy_container_bytes: header plus payload size. */
enum { GOOFY_CONTAINER_DDS = 0, GOOFY_CONTAINER_KTX = 1 };
GOOFY_HIP_API size_t goofy_container_bytes(int kind, int format, unsigned int width, unsigned int height);
GOOFY_HIP_API int goofy_container_write(int kind, int format, unsigned int width, unsigned int height,
                                        const void* blocks, size_t blocks_bytes, void* out, size_t out_capacity,
                                        size_t* written);
GOOFY_HIP_API int goofy_container_write_file(const char* path, int kind, int format, unsigned int width,
                                             unsigned int height, const void* blocks, size_t blocks_bytes);

/* ---- synthetic inputs ------------------------------------------------------
 * Seeded, integer-only texture generator (host and device give identical
 * bytes).  kind 0: three 2-D sinusoids per channel + uniform noise in
 * [-noise, noise]; kind 1: 64x64 stress regions (constant, checker,
 * saturated primaries, noise, normal map).  Alpha is 255; pitch padding is
 * left untouched. */
GOOFY_HIP_API int goofy_synth_fill_host(int kind, uint64_t seed, uint32_t texture_id, uint32_t noise,
                                        unsigned char* dst, unsigned int width, unsigned int height, size_t pitch,
                                        int threads);
GOOFY_HIP_API int goofy_synth_fill_device(int kind, uint64_t seed, uint32_t texture_id, uint32_t noise, void* d_dst,
                                          unsigned int width, unsigned int height, size_t pitch, void* stream);

#ifdef __cplusplus
}
#endif
#endif
