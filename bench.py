"""Benchmark of the MI355X Goofy encoder; prints ONE JSON line.

Headline (`value`): plain encode of one 8192x8192 RGBA8 image (seeded waves +
noise 12) already resident in HBM; one step = the DXT1 encode plus the ETC1S
encode of that image (one kernel launch each; the two independent launches
go to two streams), so MP/s counts the pixels of both formats per second of
wall time over the timed steps.  Four input buffers of 256 MiB are rotated so
that no launch reads an image the 256 MB last-level cache still holds.

Also reported:
* from a second timed region of the same steps on one stream, per-format
  kernel times (HIP events recorded by each kernel's own dispatch) and the HBM
  roofline of the slower kernel, beside a traffic probe (same addresses, no
  encode arithmetic, best of four workgroup sizes) that gives the attainable
  floor;
* both formats from one read of the 8192^2 image (the dual-format entry
  point);
* DXT5 (BC3) and ETC2 RGBA8 on the 8192^2 image with an alpha channel;
* the 16384^2 stress image with its padded pitch (default and tuned launch
  shapes);
* the decoders, the RGB565 quality floor and the quality sums on the 8192^2
  image;
* the batched workload: 4096 textures in one autotuned graph launch against
  4096 plain launches;
* the serving workload: Poisson arrivals, p50/p99 latency, and capacity
  batched against one launch per request;
* the drop-in on host buffers (PCIe included);
* the reference encoder on the host CPU (compiled from its own source, one
  stripe of block rows per thread).
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "MP/s bit-exact DXT1/ETC1S (batched+served vs plain launch); p50/p99 latency"
HBM_PEAK_GBPS = 8000.0
CLOCK_RAMP_S = 0.05  # wall time of encode load before the first timed region


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


# ---- plain 8192^2 ------------------------------------------------------------------------

def bench_plain(g, hip, W, steps: int, warmup: int, rotate: int, cfg: dict | None):
    im = W.IMAGE_8192
    nbytes = im.pitch * im.height
    out_bytes = im.blocks * 8
    log(f"plain: generating {rotate} x {im.name} on device")
    srcs = []
    stream = hip.Stream()
    for _ in range(rotate):
        s = hip.DeviceBuffer(nbytes)
        g.synth_device(im.kind, im.seed, im.texture_id, im.noise, s, im.width, im.height, im.pitch, stream)
        srcs.append(s)
    outs = {f: [hip.DeviceBuffer(out_bytes) for _ in range(rotate)] for f in (W.DXT1, W.ETC1)}
    stream.synchronize()

    # parity of the measured path against the reference's digests
    digests = json.loads((ROOT / "tests" / "golden" / "digests.json").read_text())["images"][im.name]
    parity = {}
    for f, key in ((W.DXT1, "dxt1"), (W.ETC1, "etc1")):
        g.compress_device(f, outs[f][0], srcs[0], im.width, im.height, im.pitch, stream, cfg)
        stream.synchronize()
        parity[key] = sha(outs[f][0].download()) == digests[f"{key}_sha256"]
    parity["input"] = sha(srcs[0].download()) == digests["input_sha256"]

    j = 0
    stream_b = hip.Stream()

    def step(events=None):
        """One DXT1 and one ETC1S encode.  Timed for the roofline: both on one
        stream, each with start/stop events recorded by its own dispatch.
        Timed for throughput: the two independent launches on two streams."""
        nonlocal j
        for f in (W.DXT1, W.ETC1):
            k = j % rotate
            if events is not None:
                events[f].append((hip.Event(), hip.Event()))
                g.compress_device_timed(f, outs[f][k], srcs[k], im.width, im.height, im.pitch,
                                        events[f][-1][0], events[f][-1][1], stream, cfg)
            else:
                g.compress_device(f, outs[f][k], srcs[k], im.width, im.height, im.pitch,
                                  stream if f == W.DXT1 else stream_b, cfg)
            j += 1

    log(f"plain: warmup {warmup} steps, then encodes until {CLOCK_RAMP_S * 1e3:.0f} ms of load")
    for _ in range(warmup):
        step()
    stream.synchronize()
    stream_b.synchronize()
    # The GPU reaches its full clocks only after some milliseconds of load: keep
    # encoding (the same two launches per step) until CLOCK_RAMP_S of wall time
    # has passed, however few warm-up steps were asked for.
    t_ramp = time.perf_counter()
    while time.perf_counter() - t_ramp < CLOCK_RAMP_S:
        for _ in range(8):
            step()
        stream.synchronize()
        stream_b.synchronize()
    # timed region 1 (value): wall clock around exactly `steps` steps
    log(f"plain: timing {steps} steps")
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    stream.synchronize()
    stream_b.synchronize()
    elapsed = time.perf_counter() - t0
    # timed region 2 (roofline): the same steps on one stream, per-kernel events
    events = {W.DXT1: [], W.ETC1: []}
    t1 = time.perf_counter()
    for _ in range(steps):
        step(events)
    stream.synchronize()
    elapsed_single = time.perf_counter() - t1
    kern_ms = {f: float(np.mean([a.elapsed_ms(b) for a, b in events[f]])) for f in events}
    kern_med = {f: float(np.median([a.elapsed_ms(b) for a, b in events[f]])) for f in events}
    kern_ms["single_stream_ms_per_step"] = elapsed_single / steps * 1e3
    # the same traffic without the encode arithmetic: the attainable floor,
    # best over workgroup sizes; timed like the encoders (the mean of `steps`
    # dispatch-recorded launches, the clocks already up from the encodes)
    kern_ms["probe_by_wg"] = {}
    for wg in (1, 2, 4, 8):
        probe = []
        for i in range(steps):
            a, b = hip.Event(), hip.Event()
            g.traffic_probe(outs[W.DXT1][i % rotate], srcs[i % rotate], im.width, im.height, im.pitch, a, b, stream,
                            wg_waves=wg)
            probe.append((a, b))
        stream.synchronize()
        kern_ms["probe_by_wg"][wg] = float(np.mean([a.elapsed_ms(b) for a, b in probe]))
    kern_ms["probe"] = min(kern_ms["probe_by_wg"].values())
    for b in srcs + outs[W.DXT1] + outs[W.ETC1]:
        b.free()
    stream.destroy()
    stream_b.destroy()
    return im, elapsed, kern_ms, kern_med, parity


# ---- both formats in one pass ------------------------------------------------------------------

def bench_dual(g, hip, W, launches: int = 200, rotate: int = 4):
    """goofy_hip_compress_device_dual on the 8192^2 image: one read of the image,
    DXT1 and ETC1S written (64 + 16 B per block).  Kernel time from
    dispatch-recorded events over `launches` launches on one stream, inputs
    rotated; outputs checked against the reference's digests."""
    im = W.IMAGE_8192
    digests = json.loads((ROOT / "tests" / "golden" / "digests.json").read_text())["images"][im.name]
    stream = hip.Stream()
    srcs = [hip.DeviceBuffer(im.pitch * im.height) for _ in range(rotate)]
    for s_ in srcs:
        g.synth_device(im.kind, im.seed, im.texture_id, im.noise, s_, im.width, im.height, im.pitch, stream)
    a, b = hip.DeviceBuffer(im.blocks * 8), hip.DeviceBuffer(im.blocks * 8)
    g.compress_device_dual(a, b, srcs[0], im.width, im.height, im.pitch, stream)
    stream.synchronize()
    exact = sha(a.download()) == digests["dxt1_sha256"] and sha(b.download()) == digests["etc1_sha256"]
    for i in range(200):  # warm-up
        g.compress_device_dual(a, b, srcs[i % rotate], im.width, im.height, im.pitch, stream)
    stream.synchronize()
    # wall clock over back-to-back launches, then kernel time from events
    t0 = time.perf_counter()
    for i in range(launches):
        g.compress_device_dual(a, b, srcs[i % rotate], im.width, im.height, im.pitch, stream)
    stream.synchronize()
    wall_ms = (time.perf_counter() - t0) / launches * 1e3
    ev = [(hip.Event(), hip.Event()) for _ in range(launches)]
    for i, (e0, e1) in enumerate(ev):
        g.compress_device_dual(a, b, srcs[i % rotate], im.width, im.height, im.pitch, stream, e0, e1)
    stream.synchronize()
    ms = float(np.mean([e0.elapsed_ms(e1) for e0, e1 in ev]))
    algo = im.blocks * 80
    res = {"workload": f"{im.name} RGBA8, DXT1 + ETC1S from one read of the image",
           "kernel_ms_mean": round(ms, 5), "wall_ms_per_launch": round(wall_ms, 5),
           "mps_both_formats": round(2 * im.pixels / 1e6 / (wall_ms / 1e3), 1),
           "gbps": round(algo / (ms / 1e3) / 1e9, 1), "algorithmic_bytes_per_launch": algo, "bit_exact": exact}
    for x in srcs + [a, b]:
        x.free()
    stream.destroy()
    return res


# ---- formats with alpha ---------------------------------------------------------------------------

def bench_alpha(g, hip, W, launches: int = 200, rotate: int = 4):
    """DXT5 (BC3) and ETC2 RGBA8 on the 8192^2 image with an alpha channel:
    64 B read + 16 B written per block.  Kernel time from dispatch-recorded
    events over `launches` launches per format on one stream, inputs rotated;
    outputs checked against the golden digests.  Then both formats from one
    read of the image (`dual`, 64 + 16 + 16 B per block)."""
    im = W.ALPHA_8192
    digests = json.loads((ROOT / "tests" / "golden" / "digests.json").read_text())["alpha_images"][im.name]
    stream = hip.Stream()
    srcs = [hip.DeviceBuffer(im.pitch * im.height) for _ in range(rotate)]
    for s_ in srcs:
        g.synth_device(im.kind, im.seed, im.texture_id, im.noise, s_, im.width, im.height, im.pitch, stream)
    out = hip.DeviceBuffer(im.blocks * 16)
    algo = im.blocks * 80
    res = {"workload": f"{im.name} RGBA8 (alpha: opaque, cutout and smooth 64x64 regions)",
           "algorithmic_bytes_per_launch": algo}
    for fmt, key in ((W.DXT5, "dxt5"), (W.ETC2, "etc2")):
        g.compress_device(fmt, out, srcs[0], im.width, im.height, im.pitch, stream)
        stream.synchronize()
        exact = sha(out.download()) == digests[f"{key}_sha256"]
        for i in range(200):  # warm-up
            g.compress_device(fmt, out, srcs[i % rotate], im.width, im.height, im.pitch, stream)
        ev = [(hip.Event(), hip.Event()) for _ in range(launches)]
        for i, (e0, e1) in enumerate(ev):
            g.compress_device_timed(fmt, out, srcs[i % rotate], im.width, im.height, im.pitch, e0, e1, stream)
        stream.synchronize()
        t = np.array([e0.elapsed_ms(e1) for e0, e1 in ev])
        ms = float(t.mean())
        res[key] = {"kernel_ms_mean": round(ms, 5), "kernel_ms_median": round(float(np.median(t)), 5),
                    "kernel_mps": round(im.pixels / 1e6 / (ms / 1e3), 1),
                    "gbps": round(algo / (ms / 1e3) / 1e9, 1), "bit_exact": exact}
    # both 16-byte formats from one read of the image
    out2 = hip.DeviceBuffer(im.blocks * 16)
    g.compress_device_dual_alpha(out, out2, srcs[0], im.width, im.height, im.pitch, stream)
    stream.synchronize()
    exact = (sha(out.download()) == digests["dxt5_sha256"] and sha(out2.download()) == digests["etc2_sha256"])
    for i in range(200):
        g.compress_device_dual_alpha(out, out2, srcs[i % rotate], im.width, im.height, im.pitch, stream)
    ev = [(hip.Event(), hip.Event()) for _ in range(launches)]
    for i, (e0, e1) in enumerate(ev):
        g.compress_device_dual_alpha(out, out2, srcs[i % rotate], im.width, im.height, im.pitch, stream, e0, e1)
    stream.synchronize()
    ms = float(np.mean([e0.elapsed_ms(e1) for e0, e1 in ev]))
    res["dual"] = {"kernel_ms_mean": round(ms, 5), "mps_both_formats": round(2 * im.pixels / 1e6 / (ms / 1e3), 1),
                   "gbps": round(im.blocks * 96 / (ms / 1e3) / 1e9, 1), "algorithmic_bytes_per_launch": im.blocks * 96,
                   "bit_exact": exact}
    for x in srcs + [out, out2]:
        x.free()
    stream.destroy()
    return res


# ---- branch-coverage stress image ---------------------------------------------------------------

def bench_stress(g, hip, W, launches: int = 20):
    """16384x16384 stress texture (64x64 regions: constant, checker, saturated
    primaries, noise, normal map), pitch 65792 B: both formats, kernel time from
    dispatch-recorded events, output digests checked against the reference's."""
    im = W.STRESS_16384
    digests = json.loads((ROOT / "tests" / "golden" / "digests.json").read_text())["images"][im.name]
    stream = hip.Stream()
    src = hip.DeviceBuffer(im.pitch * im.height, fill=0)
    g.synth_device(im.kind, im.seed, im.texture_id, im.noise, src, im.width, im.height, im.pitch, stream)
    out = hip.DeviceBuffer(im.blocks * 8)
    algo = W.algorithmic_bytes(im.width, im.height)
    res = {"workload": f"{im.name} RGBA8 (1 GiB of pixels), stress regions, both formats"}
    for f, key in ((W.DXT1, "dxt1"), (W.ETC1, "etc1")):
        g.compress_device(f, out, src, im.width, im.height, im.pitch, stream)
        stream.synchronize()
        exact = sha(out.download()) == digests[f"{key}_sha256"]
        ev = [(hip.Event(), hip.Event()) for _ in range(launches)]
        for a, b in ev:
            g.compress_device_timed(f, out, src, im.width, im.height, im.pitch, a, b, stream)
        stream.synchronize()
        ms = float(np.mean([a.elapsed_ms(b) for a, b in ev]))
        res[key] = {"kernel_ms_mean": round(ms, 4), "mps": round(im.pixels / 1e6 / (ms / 1e3), 1),
                    "gbps": round(algo / (ms / 1e3) / 1e9, 1), "bit_exact": exact}
        # the launch-shape tuner on this shape, then the same timing with its choice
        cfg, _ = g.tune_device(f, out, src, im.width, im.height, im.pitch, stream, reps=3)
        ev = [(hip.Event(), hip.Event()) for _ in range(launches)]
        for a, b in ev:
            g.compress_device_timed(f, out, src, im.width, im.height, im.pitch, a, b, stream, config=cfg)
        stream.synchronize()
        ms = float(np.mean([a.elapsed_ms(b) for a, b in ev]))
        res[key]["tuned"] = {"config": cfg, "kernel_ms_mean": round(ms, 4),
                             "mps": round(im.pixels / 1e6 / (ms / 1e3), 1), "gbps": round(algo / (ms / 1e3) / 1e9, 1),
                             "bit_exact": sha(out.download()) == digests[f"{key}_sha256"]}
    src.free()
    out.free()
    stream.destroy()
    return res


# ---- batched --------------------------------------------------------------------------------

def bench_batch(g, hip, W, reps: int = 7, align: int = 16, per_texture: bool = True, widths=(0, 1 << 30)):
    """The 4096-texture batch as one hipGraph replay, autotuned shape.  Rows
    are `align`-byte aligned (16: the tight pitch width * 4); with
    per_texture the same textures also run as one plain launch each, for the
    speedup and a block-for-block comparison.  widths = [lo, hi): only the
    textures of those widths."""
    ts = [t for t in W.batch_textures() if widths[0] <= t["width"] < widths[1]]
    pitches = [(t["width"] * 4 + align - 1) // align * align for t in ts]
    sizes = [p * t["height"] for p, t in zip(pitches, ts)]
    offs = np.concatenate([[0], np.cumsum([(s + 255) // 256 * 256 for s in sizes])])
    osz = [(t["width"] // 4) * (t["height"] // 4) * 8 for t in ts]
    ooffs = np.concatenate([[0], np.cumsum([(s + 255) // 256 * 256 for s in osz])])
    log(f"batch: {len(ts)} textures, {offs[-1] / 2**30:.2f} GiB in, generating on device")
    stream = hip.Stream()
    src = hip.DeviceBuffer(int(offs[-1]))
    dst = hip.DeviceBuffer(int(ooffs[-1]), fill=0)
    dst2 = hip.DeviceBuffer(int(ooffs[-1]), fill=0)
    base, obase, obase2 = src.data_ptr(), dst.data_ptr(), dst2.data_ptr()
    items, items2 = [], []
    for i, t in enumerate(ts):
        g.synth_device(t["kind"], t["seed"], t["texture_id"], t["noise"], base + int(offs[i]), t["width"], t["height"],
                       pitches[i], stream=stream)
        d = dict(src=base + int(offs[i]), width=t["width"], height=t["height"], pitch=pitches[i],
                 format=t["format"])
        items.append({**d, "dst": obase + int(ooffs[i])})
        items2.append({**d, "dst": obase2 + int(ooffs[i])})
    stream.synchronize()
    b = g.Batch(items)
    b2 = g.Batch(items2)
    info = b.info()
    log("batch: autotuning")
    chosen, tuned_ms = b.autotune(stream, reps=2)

    def timed(fn, n):
        ts_ = []
        for _ in range(n):
            e0, e1 = hip.Event(), hip.Event()
            e0.record(stream)
            fn(stream)
            e1.record(stream)
            e1.synchronize()
            ts_.append(e0.elapsed_ms(e1))
        return float(np.median(ts_))

    b.launch(stream)
    stream.synchronize()
    graph_ms = timed(b.launch, reps)
    px = info["pixels"]
    res = {
        "workload": "4096 textures, W,H log-uniform multiples of 4 in [64,2048], 50/50 DXT1/ETC1S, noise U[0,40]"
                    + ("" if align == 16 else f", rows {align}-byte aligned")
                    + ("" if widths == (0, 1 << 30) else f", only the {len(ts)} of width in [{widths[0]}, {widths[1]})"),
        "textures": info["textures"], "megapixels": round(px / 1e6, 2),
        "graph_ms": round(graph_ms, 4), "mps": round(px / 1e6 / (graph_ms / 1e3), 1),
        "gbps": round(px / 16 * 72 / (graph_ms / 1e3) / 1e9, 1), "autotuned_config": chosen,
    }
    if per_texture:
        b2.launch_per_texture(stream)
        stream.synchronize()
        plain_ms = timed(b2.launch_per_texture, max(3, reps // 2))
        stream.synchronize()
        res.update({"per_texture_launch_ms": round(plain_ms, 4),
                    "per_texture_mps": round(px / 1e6 / (plain_ms / 1e3), 1),
                    "speedup_vs_per_texture": round(plain_ms / graph_ms, 2),
                    "bit_exact_vs_per_texture": bool(np.array_equal(dst.download(), dst2.download()))})
    b.close()
    b2.close()
    for x in (src, dst, dst2):
        x.free()
    stream.destroy()
    return res


# ---- the kernels either side of the encoder ------------------------------------------------------

def bench_around(g, hip, W, reps: int = 20):
    """Decoders, the harness's RGB565 quality floor and the quality sums on the
    8192^2 image: median time between events recorded on the kernels' stream
    (decoders, floor: one launch each) or around the whole call (quality:
    zeroing kernel, sum kernel, read-back), and GB/s of algorithmic bytes --
    decode 8 B in + 64 B out per block, floor 4 B in + 4 B out per pixel,
    round-trip quality 8 + 64 B read per block, image quality 8 B read per pixel.
    The DXT5 / ETC2 RGBA8 decoders and round trips run on the image with an
    alpha channel (16 B blocks: 80 B per block)."""
    im = W.IMAGE_8192
    stream = hip.Stream()
    src = hip.DeviceBuffer(im.pitch * im.height)
    g.synth_device(im.kind, im.seed, im.texture_id, im.noise, src, im.width, im.height, im.pitch, stream)
    blocks = hip.DeviceBuffer(im.blocks * 8)
    img = hip.DeviceBuffer(im.pitch * im.height)
    res = {"workload": f"{im.name} RGBA8"}

    def timed(fn, n=reps):
        ts_ = []
        for _ in range(n):
            e0, e1 = hip.Event(), hip.Event()
            e0.record(stream)
            fn()
            e1.record(stream)
            e1.synchronize()
            ts_.append(e0.elapsed_ms(e1))
        return float(np.median(ts_))

    def entry(ms, nbytes):
        return {"ms": round(ms, 4), "gbps": round(nbytes / (ms / 1e3) / 1e9, 1),
                "mps": round(im.pixels / 1e6 / (ms / 1e3), 1)}

    for f, key in ((W.DXT1, "dxt1"), (W.ETC1, "etc1")):
        g.compress_device(f, blocks, src, im.width, im.height, im.pitch, stream)
        g.decode_device(f, img, blocks, im.width, im.height, im.pitch, stream)
        res[f"decode_{key}"] = entry(timed(lambda: g.decode_device(f, img, blocks, im.width, im.height, im.pitch,
                                                                   stream)), im.blocks * 72)
        q = {}
        ms = timed(lambda: q.update(g.quality_device(f, blocks, src, im.width, im.height, im.pitch, stream)), 5)
        res[f"quality_{key}"] = {**entry(ms, im.blocks * 72), "psnr_rgb": round(q["psnr_rgb"], 3),
                                 "psnr_y": round(q["psnr_y"], 3)}
    ia = W.ALPHA_8192
    asrc = hip.DeviceBuffer(ia.pitch * ia.height)
    g.synth_device(ia.kind, ia.seed, ia.texture_id, ia.noise, asrc, ia.width, ia.height, ia.pitch, stream)
    blocks16 = hip.DeviceBuffer(ia.blocks * 16)
    for f, key in ((W.DXT5, "dxt5"), (W.ETC2, "etc2")):
        g.compress_device(f, blocks16, asrc, ia.width, ia.height, ia.pitch, stream)
        g.decode_device(f, img, blocks16, ia.width, ia.height, ia.pitch, stream)
        res[f"decode_{key}"] = entry(timed(lambda: g.decode_device(f, img, blocks16, ia.width, ia.height, ia.pitch,
                                                                   stream)), ia.blocks * 80)
        q = {}
        ms = timed(lambda: q.update(g.quality_device(f, blocks16, asrc, ia.width, ia.height, ia.pitch, stream)), 5)
        res[f"quality_{key}"] = {**entry(ms, ia.blocks * 80), "psnr_rgb": round(q["psnr_rgb"], 3),
                                 "psnr_a": round(q["psnr_a"], 3)}
    asrc.free()
    blocks16.free()
    g.rgb565_baseline_device(img, src, im.width, im.height, im.pitch, im.pitch, stream)
    res["rgb565_floor"] = entry(timed(lambda: g.rgb565_baseline_device(img, src, im.width, im.height, im.pitch,
                                                                       im.pitch, stream)), im.pixels * 8)
    q = {}
    ms = timed(lambda: q.update(g.quality_images_device(src, img, im.width, im.height, im.pitch, im.pitch, stream)),
               5)
    res["quality_floor"] = {**entry(ms, im.pixels * 8), "psnr_rgb": round(q["psnr_rgb"], 3)}
    for x in (src, blocks, img):
        x.free()
    stream.destroy()
    return res


# ---- drop-in with host buffers (PCIe-inclusive) -------------------------------------------------

def bench_dropin(g, W, reps: int = 5):
    """goofy::compressDXT1/ETC1 on host memory: staging, kernel and copy-back, as a caller
    of the reference API sees it.  Best of `reps` per format and size."""
    res = {}
    for im in (W.IMAGE_512, W.IMAGE_8192):
        img = g.synth_host(im.kind, im.seed, im.texture_id, im.noise, im.width, im.height, im.pitch)
        out = np.empty(im.blocks * 8, np.uint8)
        for f, key, fn in ((W.DXT1, "dxt1", g.compressDXT1), (W.ETC1, "etc1", g.compressETC1)):
            assert fn(out, img, im.width, im.height, im.pitch) == 0
            best = 1e9
            for _ in range(reps if im is W.IMAGE_8192 else 4 * reps):
                t0 = time.perf_counter()
                fn(out, img, im.width, im.height, im.pitch)
                best = min(best, time.perf_counter() - t0)
            res[f"{im.name}_{key}"] = {"ms": round(best * 1e3, 3), "mps": round(im.pixels / 1e6 / best, 1),
                                        "pcie_gbps": round((im.pitch * im.height + im.blocks * 8) / best / 1e9, 2)}
        # both formats from one transfer of the image (goofy_hip_compress_dual)
        pair = (out, np.empty(im.blocks * 8, np.uint8))
        g.compress_dual(img, im.width, im.height, im.pitch, out=pair)
        best = 1e9
        for _ in range(reps if im is W.IMAGE_8192 else 4 * reps):
            t0 = time.perf_counter()
            g.compress_dual(img, im.width, im.height, im.pitch, out=pair)
            best = min(best, time.perf_counter() - t0)
        res[f"{im.name}_dual"] = {"ms": round(best * 1e3, 3), "mps": round(im.pixels / 1e6 / best, 1),
                                  "pcie_gbps": round((im.pitch * im.height + im.blocks * 16) / best / 1e9, 2)}
    return res


# ---- CPU baseline ------------------------------------------------------------------------------

def cpu_share() -> tuple[int, int, int]:
    """(threads to use, CPUs this process may run on, CPUs the host has).  A
    GPU box allots each GPU a share of its host CPUs (16, also exported as
    OMP_NUM_THREADS there); the baseline runs one thread per CPU of that
    share, and on every CPU where the process's affinity allows fewer."""
    visible = os.cpu_count() or 1
    allowed = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else visible
    try:
        share = int(os.environ.get("OMP_NUM_THREADS", "16"))
    except ValueError:
        share = 16
    return max(1, min(allowed, share if share > 0 else 16)), allowed, visible


def cpu_baseline(g, W, budget_s: float = 3.0):
    from oracle.oracle import Oracle, Reference, aligned_empty, have_reference
    threads, allowed, visible = cpu_share()
    im = W.IMAGE_8192
    img = g.synth_host(im.kind, im.seed, im.texture_id, im.noise, im.width, im.height, im.pitch)
    kind = "reference" if have_reference() else "port"
    if kind == "reference":
        ref = Reference()
        outs = {}

        def enc(f, img_, w, h, p, t):
            if w not in outs:
                outs[w] = aligned_empty((w // 4) * (h // 4) * 8)
            ref.compress_mt(f, img_, w, h, p, t, out=outs[w])
    else:
        ora = Oracle()

        def enc(f, img_, w, h, p, t):
            ora.compress_blocks(f, img_, w, h, p, threads=t)
    times = {0: [], 1: []}
    t_start = time.perf_counter()
    while time.perf_counter() - t_start < budget_s or min(len(v) for v in times.values()) < 2:
        for f in (0, 1):
            t0 = time.perf_counter()
            enc(f, img, im.width, im.height, im.pitch, threads)
            times[f].append(time.perf_counter() - t0)
        if min(len(v) for v in times.values()) >= 64:
            break
    best = {f: min(v) for f, v in times.items()}
    reps = len(times[0])
    # the reference's own single-thread case: 512^2, best of 128 (Src/main.cpp:43,653-664)
    s = W.IMAGE_512
    small = g.synth_host(s.kind, s.seed, s.texture_id, s.noise, s.width, s.height, s.pitch)
    st = {}
    for f in (0, 1):
        bt = 1e9
        for _ in range(128):
            t0 = time.perf_counter()
            enc(f, small, s.width, s.height, s.pitch, 1)
            bt = min(bt, time.perf_counter() - t0)
        st["dxt1" if f == 0 else "etc1"] = round(s.pixels / 1e6 / bt, 1)
    mps = 2 * im.pixels / 1e6 / (best[0] + best[1])
    return {
        "value": round(mps, 1), "unit": "MP/s", "cores": threads, "kind": kind,
        "host_cpus": visible, "affinity_cpus": allowed,
        "threads_rule": "one thread per CPU of the GPU's host-CPU share (OMP_NUM_THREADS, 16 per GPU on the box)",
        "per_thread_mps": round(mps / threads, 1),
        "sample": f"{im.name} DXT1+ETC1S, {reps} encodes per format, best of, one stripe of block rows per thread",
        "per_format_mps": {"dxt1": round(im.pixels / 1e6 / best[0], 1), "etc1": round(im.pixels / 1e6 / best[1], 1)},
        "single_thread_512_mps": st,
    }


def load_traffic():
    p = ROOT / "profiles" / "hbm_traffic.json"
    if not p.exists():
        return None
    try:
        return json.loads(p.read_text())
    except Exception:
        return None


def profile_mode(g, hip, W, launches: int):
    """Short run for rocprofv3, all on one stream: each plain kernel `launches`
    times (default launch shape), the traffic probe `launches` times per
    workgroup size (1, 2, 4, 8 waves), the dual-format kernel `launches` times,
    the DXT5 / ETC2 RGBA8 kernels and their one-read pair `launches` times each
    on the image with an alpha channel, the four decoders and the four round-trip
    quality calls `launches` times each on that image, and a device-to-device copy
    of the 256 MiB image (the runtime's copy kernel) used to calibrate the HBM byte
    counters.  Encoder inputs rotate over four 256 MiB buffers."""
    im = W.IMAGE_8192
    nbytes = im.pitch * im.height
    srcs = [hip.DeviceBuffer(nbytes) for _ in range(4)]
    for s in srcs:
        g.synth_device(im.kind, im.seed, im.texture_id, im.noise, s, im.width, im.height, im.pitch)
    out = hip.DeviceBuffer(im.blocks * 8)
    copy_dst = hip.DeviceBuffer(nbytes)
    hip.synchronize()
    for i in range(launches):
        copy_dst.copy_from(srcs[i % 4])  # calibration: reads and writes 256 MiB
    for f in (W.DXT1, W.ETC1):
        for i in range(launches):
            g.compress_device(f, out, srcs[i % 4], im.width, im.height, im.pitch)
    for wg in (1, 2, 4, 8):
        for i in range(launches):
            g.traffic_probe(out, srcs[i % 4], im.width, im.height, im.pitch, wg_waves=wg)
    out2 = hip.DeviceBuffer(im.blocks * 8)
    for i in range(launches):
        g.compress_device_dual(out, out2, srcs[i % 4], im.width, im.height, im.pitch)
    hip.synchronize()
    ia = W.ALPHA_8192
    for s in srcs:
        g.synth_device(ia.kind, ia.seed, ia.texture_id, ia.noise, s, ia.width, ia.height, ia.pitch)
    out16 = hip.DeviceBuffer(ia.blocks * 16)
    for f in (W.DXT5, W.ETC2):
        for i in range(launches):
            g.compress_device(f, out16, srcs[i % 4], ia.width, ia.height, ia.pitch)
    out16b = hip.DeviceBuffer(ia.blocks * 16)
    for i in range(launches):
        g.compress_device_dual_alpha(out16, out16b, srcs[i % 4], ia.width, ia.height, ia.pitch)
    hip.synchronize()
    # the decoders and the round-trip quality sums, all four formats on the image with alpha
    # (blocks of srcs[0]): `launches` decodes per format, then `launches` quality calls
    src = srcs[0]
    sets = ((W.DXT1, out), (W.ETC1, out2), (W.DXT5, out16), (W.ETC2, out16b))
    for f, blk in sets:
        g.compress_device(f, blk, src, ia.width, ia.height, ia.pitch)
    img = copy_dst
    for f, blk in sets:
        for _ in range(launches):
            g.decode_device(f, img, blk, ia.width, ia.height, ia.pitch)
    for f, blk in sets:
        for _ in range(launches):
            g.quality_device(f, blk, src, ia.width, ia.height, ia.pitch)
    hip.synchronize()
    log(f"profile mode done: {launches} launches per kernel")


def sweep_mode(g, hip, W, launches: int = 40):
    """Launch-shape sweep of the plain kernels: median kernel time
    (dispatch-recorded events) per blocks/lane, waves/workgroup, streaming and
    tile order, on the 8192^2 image (four rotated inputs) and on the 16384^2
    stress image with its 65792-byte pitch (streaming only)."""
    def table(im, nbuf, nts):
        srcs = [hip.DeviceBuffer(im.pitch * im.height) for _ in range(nbuf)]
        for s in srcs:
            g.synth_device(im.kind, im.seed, im.texture_id, im.noise, s, im.width, im.height, im.pitch)
        out = hip.DeviceBuffer(im.blocks * 8)
        hip.synchronize()
        algo = W.algorithmic_bytes(im.width, im.height)
        print(f"\n{im.name}, pitch {im.pitch} B, {nbuf} input(s)\n")
        print("| format | blocks/lane | waves/WG | streaming | tile order | median us | GB/s |")
        print("|---|---|---|---|---|---|---|")
        for fmt in (W.DXT1, W.ETC1):
            for bpl in (1, 2, 4):
                for wg in (1, 2, 4, 8):
                    for nt in nts:
                        for order in (0, 1):
                            cfg = dict(blocks_per_lane=bpl, wg_waves=wg, waves_per_cu=0, nontemporal=nt,
                                       tile_order=order)
                            for i in range(4):
                                g.compress_device(fmt, out, srcs[i % nbuf], im.width, im.height, im.pitch,
                                                  config=cfg)
                            ev = [(hip.Event(), hip.Event()) for _ in range(launches)]
                            for i, (a, b) in enumerate(ev):
                                g.compress_device_timed(fmt, out, srcs[i % nbuf], im.width, im.height, im.pitch, a,
                                                        b, config=cfg)
                            hip.synchronize()
                            med = float(np.median([a.elapsed_ms(b) for a, b in ev]))
                            print(f"| {'DXT1' if fmt == W.DXT1 else 'ETC1S'} | {bpl} | {wg} | {nt} | {order} | "
                                  f"{med * 1e3:.2f} | {algo / med / 1e6:.0f} |", flush=True)
        for x in srcs + [out]:
            x.free()

    table(W.IMAGE_8192, 4, (0, 1))
    table(W.STRESS_16384, 1, (1,))


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=200,
                    help="untimed steps first; the GPU needs about 10 ms of load to reach full clocks")
    ap.add_argument("--rotate", type=int, default=4)
    ap.add_argument("--skip-batch", action="store_true")
    ap.add_argument("--skip-serve", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-dropin", action="store_true")
    ap.add_argument("--skip-stress", action="store_true")
    ap.add_argument("--skip-around", action="store_true")
    ap.add_argument("--skip-dual", action="store_true")
    ap.add_argument("--skip-alpha", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=10.0)
    ap.add_argument("--profile", type=int, default=0, help="rocprof mode: N launches per kernel, no JSON")
    ap.add_argument("--sweep", action="store_true", help="launch-shape sweep table, no JSON")
    ap.add_argument("--batch-align", action="store_true",
                    help="the batch at tight, 128-B and 256-B aligned row pitches, then its narrow and wide "
                         "textures alone; one JSON object each")
    ap.add_argument("--bpl", type=int, default=0)
    ap.add_argument("--wg", type=int, default=0)
    ap.add_argument("--nt", type=int, default=-1)
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the encoder needs no collective, and the multi-GPU "
              "configuration is not part of this build (run with --gpus 1)", file=sys.stderr)
        return 2

    import goofy_amd as g
    from goofy_amd import hip
    from goofy_amd import workloads as W

    assert hip.device_count() > 0, "bench.py needs a GPU"
    if args.profile:
        profile_mode(g, hip, W, args.profile)
        return 0
    if args.sweep:
        sweep_mode(g, hip, W)
        return 0
    if args.batch_align:
        for align in (16, 128, 256, 16):
            print(json.dumps({"row_alignment": align, **bench_batch(g, hip, W, reps=15, align=align,
                                                                    per_texture=False)}), flush=True)
        for widths in ((0, 256), (256, 1 << 30), (1024, 1 << 30)):
            print(json.dumps({"row_alignment": 16, **bench_batch(g, hip, W, reps=15, per_texture=False,
                                                                 widths=widths)}), flush=True)
        return 0

    cfg = None
    if args.bpl or args.wg or args.nt >= 0:
        cfg = dict(blocks_per_lane=args.bpl or 1, wg_waves=args.wg or 8, waves_per_cu=0,
                   nontemporal=1 if args.nt < 0 else args.nt)

    im, elapsed, kern_ms, kern_med, parity = bench_plain(g, hip, W, args.steps, args.warmup, args.rotate, cfg)
    ms_per_step = elapsed / args.steps * 1e3
    value = 2 * im.pixels / 1e6 / (ms_per_step / 1e3)
    algo = W.algorithmic_bytes(im.width, im.height)
    dom = max((W.DXT1, W.ETC1), key=kern_ms.get)
    achieved = algo / (kern_ms[dom] / 1e3) / 1e9
    attainable = algo / (kern_ms["probe"] / 1e3) / 1e9
    traffic = load_traffic()
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBPS, 4),
                "traffic": (traffic or {}).get("bytes_per_launch", {}).get("dxt1" if dom == W.DXT1 else "etc1"),
                "kernel": "encode ETC1S" if dom == W.ETC1 else "encode DXT1",
                "algorithmic_bytes_per_launch": algo,
                "attainable_gbps": round(attainable, 1), "frac_of_attainable": round(achieved / attainable, 4),
                "probe_gbps_by_wg_waves": {str(k): round(algo / (v / 1e3) / 1e9, 1)
                                           for k, v in kern_ms["probe_by_wg"].items()}}
    formats = {}
    for f, key in ((W.DXT1, "dxt1"), (W.ETC1, "etc1")):
        formats[key] = {"kernel_ms_mean": round(kern_ms[f], 5), "kernel_ms_median": round(kern_med[f], 5),
                        "kernel_mps": round(im.pixels / 1e6 / (kern_ms[f] / 1e3), 1),
                        "kernel_gbps": round(algo / (kern_ms[f] / 1e3) / 1e9, 1)}
    formats["single_stream_ms_per_step"] = round(kern_ms["single_stream_ms_per_step"], 5)
    result = {
        "metric": METRIC, "value": round(value, 1), "unit": "MP/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": {"workload": f"plain encode, {im.name} RGBA8 stride {im.pitch} B, waves+noise12 seed "
                               f"{im.seed:#x}; step = DXT1 + ETC1S launch (two streams), inputs HBM-resident, "
                               "4-buffer rotation",
                   "image": f"{im.width}x{im.height}", "formats": ["DXT1", "ETC1S"], "parallelism": "single GPU"},
        "roofline": roofline, "formats": formats, "parity": parity,
    }
    if not args.skip_dual:
        log("dual-format encode 8192^2")
        result["dual"] = bench_dual(g, hip, W)
    if not args.skip_alpha:
        log("DXT5 / ETC2 RGBA8 8192^2")
        result["alpha"] = bench_alpha(g, hip, W)
    if not args.skip_stress:
        log("stress 16384^2")
        result["stress"] = bench_stress(g, hip, W)
    if not args.skip_around:
        log("decoders, quality floor, quality")
        result["around"] = bench_around(g, hip, W)
    if not args.skip_batch:
        result["batched"] = bench_batch(g, hip, W)
    if not args.skip_serve:
        from goofy_amd import serving
        result["served"] = serving.bench_serving(g, hip, W, log=log, duration_s=args.serve_seconds)
    if not args.skip_dropin:
        log("drop-in on host buffers")
        result["dropin_host"] = bench_dropin(g, W)
    if not args.skip_cpu:
        log("cpu baseline")
        result["cpu_baseline"] = cpu_baseline(g, W)
    if traffic:
        result["roofline"]["traffic_source"] = traffic.get("source")
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
