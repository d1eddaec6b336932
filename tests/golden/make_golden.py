"""Write the golden vectors under tests/golden/ from the reference encoder.

The reference (GoofyTC/goofy_tc.h) is compiled from its own source into
oracle/_ref/libgoofy_ref.so by oracle/Makefile; this script feeds it inputs
and records what it outputs:

  blocks.npz   -- special and random 4x4 blocks (64 B each) with the
                  reference's DXT1 and ETC1S outputs (8 B each);
  img512.npz   -- the full DXT1 / ETC1S outputs of the 512x512 waves image;
  real_images.json -- for eleven of the reference's own test images
                  (test-data/*.png named in REAL_IMAGES, copied under
                  tests/golden/images/ as data), from the reference harness
                  (Src/main.cpp, compiled into oracle/_ref/libgoofy_harness.so):
                  SHA-256 of the RGBA8 input as its loader returns it and of
                  both encodings; the harness's getMsePsnr figures of the round
                  trip through its decoders; the SHA-256 of its RGB565 quality
                  floor and that floor's getMsePsnr figures.  Under
                  "harness_table": the per-image PSNR table the harness prints
                  (Src/main.cpp:949-979) for Goofy DXT1 / ETC1 and the floor,
                  over every image of its long-run list (:991-1039) present in
                  the reference tree (read from there when this script runs);
  digests.json -- SHA-256 of inputs and outputs of the 512^2, 8192^2 and the
                  pitch-padded 16384^2 stress images, and of a 16-texture
                  batch with widths that are multiples of 4 but not of 16
                  (encoded by the reference on rows padded to a multiple of
                  16 pixels, then cropped block-wise); under "alpha_images",
                  the DXT5 / ETC2 RGBA8 encodings of the images with an alpha
                  channel (the reference has no alpha encoder: their alpha
                  blocks come from the oracle, whose rule tests/test_alpha.py
                  pins through the reference's decoders, and their colour
                  halves are checked against the reference's DXT1 / ETC1S
                  here before the digests are written).

Inputs come from the library's seeded integer generator; the input digests
pin that generator too.  Run: python tests/golden/make_golden.py [real|alpha]
"""
from __future__ import annotations

import hashlib
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import goofy_amd  # noqa: E402
from goofy_amd import workloads as W  # noqa: E402
from oracle.oracle import Harness, Oracle, Reference, aligned_empty  # noqa: E402

OUT = Path(__file__).resolve().parent


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def special_blocks(rng: np.random.Generator) -> np.ndarray:
    blocks = []

    def solid(r, g, b, a=255):
        return np.tile(np.array([r, g, b, a], np.uint8), 16)

    for c in [(0, 0, 0), (255, 255, 255), (255, 0, 0), (0, 255, 0), (0, 0, 255), (1, 1, 1), (7, 8, 9),
              (8, 8, 8), (9, 9, 9), (127, 128, 129), (254, 254, 254), (16, 17, 15)]:
        blocks.append(solid(*c))
    for _ in range(64):  # random constant colours, random alpha
        blocks.append(solid(*rng.integers(0, 256, 3), int(rng.integers(0, 256))))
    # grey two-level blocks covering every luma range 0..255 and several offsets
    for r in range(256):
        for lo in {0, (255 - r) // 2, 255 - r}:
            px = np.full((16, 4), 255, np.uint8)
            pat = rng.integers(0, 2, 16)
            pat[0], pat[1] = 0, 1
            v = np.where(pat == 1, lo + r, lo).astype(np.uint8)
            px[:, 0] = px[:, 1] = px[:, 2] = v
            blocks.append(px.reshape(64))
    # two-colour checkers
    for _ in range(128):
        a, b = rng.integers(0, 256, 3), rng.integers(0, 256, 3)
        px = np.zeros((4, 4, 4), np.uint8)
        px[..., 3] = 255
        for y in range(4):
            for x in range(4):
                px[y, x, :3] = a if (x + y) % 2 == 0 else b
        blocks.append(px.reshape(64))
    # saturated primaries
    for _ in range(128):
        px = (rng.integers(0, 2, (16, 4)) * 255).astype(np.uint8)
        px[:, 3] = 255
        blocks.append(px.reshape(64))
    # one outlier pixel in a flat block, at every position
    for p in range(16):
        px = np.tile(np.array([100, 120, 140, 255], np.uint8), (16, 1))
        px[p, :3] = rng.integers(0, 256, 3)
        blocks.append(px.reshape(64))
    # alpha noise over identical RGB (alpha must not change the output)
    base = rng.integers(0, 256, (16, 4)).astype(np.uint8)
    for _ in range(16):
        px = base.copy()
        px[:, 3] = rng.integers(0, 256, 16)
        blocks.append(px.reshape(64))
    # mid == qt and near-threshold cases: lumas spread just around mid
    for mid in range(0, 256, 3):
        px = np.full((16, 4), 255, np.uint8)
        v = np.clip(mid + rng.integers(-6, 7, 16), 0, 255).astype(np.uint8)
        px[:, 0] = px[:, 1] = px[:, 2] = v
        blocks.append(px.reshape(64))
    return np.stack(blocks)


def random_blocks(rng: np.random.Generator, n: int) -> np.ndarray:
    out = []
    out.append(rng.integers(0, 256, (n // 4, 64)).astype(np.uint8))
    base = rng.integers(0, 256, (n // 4, 1))
    spread = rng.integers(0, 48, (n // 4, 1))
    out.append(np.clip(base + rng.integers(-48, 49, (n // 4, 64)) % (spread + 1), 0, 255).astype(np.uint8))
    # per-channel ramps
    t = np.linspace(0, 1, 16)[None, :, None]
    a = rng.integers(0, 256, (n // 4, 1, 4))
    b = rng.integers(0, 256, (n // 4, 1, 4))
    out.append(np.clip(a + (b - a) * t + rng.integers(-3, 4, (n // 4, 16, 4)), 0, 255).astype(np.uint8).reshape(-1, 64))
    # small per-pixel noise around a colour (gradients of natural images)
    c = rng.integers(20, 236, (n // 4, 1, 4))
    out.append(np.clip(c + rng.integers(-20, 21, (n // 4, 16, 4)), 0, 255).astype(np.uint8).reshape(-1, 64))
    return np.concatenate(out)


def encode_image_ref(ref: Reference, fmt: int, img: np.ndarray, width: int, height: int, pitch: int) -> np.ndarray:
    if width % 16 == 0:
        rc, out = ref.compress(fmt, img, width, height, pitch)
        assert rc == 0
        return out
    # widths that are multiples of 4 only: pad each row to a multiple of 16
    # pixels (padding blocks are independent of the real ones), encode, crop.
    wp = (width + 15) // 16 * 16
    padded = aligned_empty(wp * 4 * height)
    padded[:] = 0
    pv = padded.reshape(height, wp * 4)
    pv[:, : width * 4] = img.reshape(height, pitch)[:, : width * 4]
    rc, out = ref.compress(fmt, padded, wp, height, wp * 4)
    assert rc == 0
    return out.reshape(height // 4, wp // 4, 8)[:, : width // 4].reshape(-1).copy()


def main() -> int:
    ref = Reference()
    rng = np.random.default_rng(20240601)
    blocks = np.concatenate([special_blocks(rng), random_blocks(rng, 2048)])
    dxt1 = ref.encode_blocks(W.DXT1, blocks)
    etc1 = ref.encode_blocks(W.ETC1, blocks)
    np.savez_compressed(OUT / "blocks.npz", blocks=blocks, dxt1=dxt1, etc1=etc1)
    print(f"blocks.npz: {blocks.shape[0]} blocks")

    digests: dict = {"images": {}, "batch": []}
    for im in (W.IMAGE_512, W.IMAGE_8192, W.STRESS_16384):
        img = goofy_amd.synth_host(im.kind, im.seed, im.texture_id, im.noise, im.width, im.height, im.pitch)
        entry = {"width": im.width, "height": im.height, "pitch": im.pitch, "input_sha256": sha(img)}
        for fmt, name in ((W.DXT1, "dxt1"), (W.ETC1, "etc1")):
            out = encode_image_ref(ref, fmt, img, im.width, im.height, im.pitch)
            entry[f"{name}_sha256"] = sha(out)
            if im is W.IMAGE_512:
                entry[f"{name}_out"] = out
        if im is W.IMAGE_512:
            np.savez_compressed(OUT / "img512.npz", dxt1=entry.pop("dxt1_out"), etc1=entry.pop("etc1_out"))
        digests["images"][im.name] = entry
        print(im.name, entry["input_sha256"][:16])
        del img

    for t in W.batch_textures(n=16, lo=64, hi=512):
        img = goofy_amd.synth_host(t["kind"], t["seed"], t["texture_id"], t["noise"], t["width"], t["height"])
        out = encode_image_ref(ref, t["format"], img, t["width"], t["height"], t["width"] * 4)
        digests["batch"].append({**t, "input_sha256": sha(img), "output_sha256": sha(out)})
    (OUT / "digests.json").write_text(json.dumps(digests, indent=1) + "\n")
    print("digests.json written")

    write_real_images(ref)
    return 0


def write_alpha_digests(ref: Reference) -> None:
    ora = Oracle()
    path = OUT / "digests.json"
    digests = json.loads(path.read_text())
    digests["alpha_images"] = {}
    for im in (W.ALPHA_512, W.ALPHA_8192):
        img = goofy_amd.synth_host(im.kind, im.seed, im.texture_id, im.noise, im.width, im.height, im.pitch)
        entry = {"width": im.width, "height": im.height, "pitch": im.pitch, "input_sha256": sha(img)}
        for fmt, cfmt, name in ((W.DXT5, W.DXT1, "dxt5"), (W.ETC2, W.ETC1, "etc2")):
            out = ora.compress_blocks(fmt, img, im.width, im.height, im.pitch)
            colour = encode_image_ref(ref, cfmt, img, im.width, im.height, im.pitch)
            assert np.array_equal(out.reshape(-1, 16)[:, 8:].reshape(-1), colour), (im.name, name)
            entry[f"{name}_sha256"] = sha(out)
        digests["alpha_images"][im.name] = entry
        print(im.name, entry["input_sha256"][:16])
    path.write_text(json.dumps(digests, indent=1) + "\n")


# Images stored under tests/golden/images/ (GPU tests read them there).
REAL_IMAGES = ("parrot_red.png", "kodim23.png", "patterns.png", "baboon.png", "monarch.png", "roblox01.png",
               "roblox05.png", "kodim03.png", "kodim20.png", "peppers.png", "lena.png")
# The harness's long-run image list (Src/main.cpp:991-1039), in its order.
HARNESS_LIST = (["patterns", "pbr_bricks_albedo", "pbr_ground_albedo", "pbr_stones_albedo", "pbr_stones_normal",
                 "pbr_head_albedo", "parrot_red", "baboon", "lena", "monarch", "peppers", "sail", "tulips"] +
                [f"kodim{i:02d}" for i in range(1, 25)] + [f"roblox{i:02d}" for i in range(1, 7)])
REFERENCE_TEST_DATA = Path("/root/reference/test-data")


def harness_table(ref: Reference, harness: Harness) -> dict:
    """psnrMin / psnrRGB / psnrY of Goofy DXT1, Goofy ETC1 and the RGB565
    floor per image, as the harness computes them, for the list's images
    present in the reference tree; plus their plain means."""
    rows = {}
    for stem in HARNESS_LIST:
        path = REFERENCE_TEST_DATA / f"{stem}.png"
        if not path.exists():
            continue
        img, w, h = harness.load_png(path)
        row = {}
        for fmt, key in ((W.DXT1, "dxt1"), (W.ETC1, "etc1")):
            out = encode_image_ref(ref, fmt, img, w, h, w * 4)
            q = harness.mse_psnr(img, harness.decompress(fmt, out, w, h), w, h)
            row[key] = {k: q[k] for k in ("psnr_min", "psnr_rgb", "psnr_y")}
        q = harness.mse_psnr(img, harness.rgb565_baseline(img, w, h), w, h)
        row["floor"] = {k: q[k] for k in ("psnr_min", "psnr_rgb", "psnr_y")}
        rows[stem] = row
    means = {key: {k: float(np.mean([r[key][k] for r in rows.values()])) for k in ("psnr_min", "psnr_rgb", "psnr_y")}
             for key in ("dxt1", "etc1", "floor")}
    missing = [s for s in HARNESS_LIST if s not in rows]
    return {"images": rows, "mean_over_present": means, "missing_from_reference_tree": missing}


def write_real_images(ref: Reference) -> None:
    harness = Harness()
    real = {}
    for name in REAL_IMAGES:
        img, w, h = harness.load_png(OUT / "images" / name)
        entry = {"width": w, "height": h, "input_sha256": sha(img)}
        for fmt, key in ((W.DXT1, "dxt1"), (W.ETC1, "etc1")):
            out = encode_image_ref(ref, fmt, img, w, h, w * 4)
            entry[f"{key}_sha256"] = sha(out)
            dec = harness.decompress(fmt, out, w, h)
            entry[f"{key}_psnr"] = harness_psnr(img, dec)
            entry[f"{key}_quality"] = harness.mse_psnr(img, dec, w, h)
        base = harness.rgb565_baseline(img, w, h)
        entry["baseline_sha256"] = sha(base)
        entry["baseline_quality"] = harness.mse_psnr(img, base, w, h)
        real[name] = entry
        print(name, w, h, {k: round(v["psnr_rgb"], 3) for k, v in entry.items() if k.endswith("quality")})
    real["harness_table"] = harness_table(ref, harness)
    print("mean over present images:", real["harness_table"]["mean_over_present"])
    (OUT / "real_images.json").write_text(json.dumps(real, indent=1) + "\n")


def harness_psnr(src: np.ndarray, dec: np.ndarray) -> dict:
    """getMsePsnr (Src/main.cpp:403-469): pixel read as a little-endian word,
    "r" = bits 16..23, "b" = bits 0..7; RGB PSNR with peak 768."""
    a = src.reshape(-1, 4).astype(np.float64)
    b = dec.reshape(-1, 4).astype(np.float64)
    n = a.shape[0]
    se = ((a - b) ** 2).sum(0)
    y1 = 0.21 * a[:, 2] + 0.72 * a[:, 1] + 0.07 * a[:, 0]
    y2 = 0.21 * b[:, 2] + 0.72 * b[:, 1] + 0.07 * b[:, 0]
    mse_y = float(((y2 - y1) ** 2).sum() / n)
    psnr = lambda m, p=255.0: None if m < 1e-7 else 10 * np.log10(p * p / m)  # noqa: E731
    return {"sse_byte": [int(x) for x in se], "psnr_rgb": psnr(se[:3].sum() / n, 768.0), "psnr_y": psnr(mse_y)}


if __name__ == "__main__":
    if sys.argv[1:] == ["real"]:  # only real_images.json
        write_real_images(Reference())
        sys.exit(0)
    if sys.argv[1:] == ["alpha"]:  # only the "alpha_images" digests
        write_alpha_digests(Reference())
        sys.exit(0)
    rc = main()
    write_alpha_digests(Reference())
    sys.exit(rc)
