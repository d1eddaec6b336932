"""Serving: continuous batching of encode requests under an HBM budget.

Thin ctypes layer over the server in libgoofy_hip.so (goofy_server_* and
goofy_serve_simulate in include/goofy_hip.h), plus the serving workload used
by bench.py: Poisson arrivals of device-resident textures, p50/p99 latency
and sustained MP/s, and a saturation run that compares batched launches with
one plain launch per request.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import LaunchConfig, TextureDesc, _check, _lib, _ptr

_vp = ctypes.c_void_p
_u64 = ctypes.c_uint64


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_size_t), ("arena_bytes", ctypes.c_size_t),
                ("streams", ctypes.c_int), ("max_batch_textures", ctypes.c_int),
                ("max_pending", ctypes.c_int), ("batching", ctypes.c_int), ("launch", LaunchConfig)]


class Request(ctypes.Structure):
    _fields_ = [("src", _vp), ("width", ctypes.c_uint), ("height", ctypes.c_uint), ("pitch", ctypes.c_size_t),
                ("format", ctypes.c_int), ("user_tag", _u64), ("arrival_us", ctypes.c_double),
                ("deadline_us", ctypes.c_double)]


class Completion(ctypes.Structure):
    _fields_ = [("ticket", _u64), ("user_tag", _u64), ("status", ctypes.c_int), ("dst", _vp),
                ("dst_bytes", ctypes.c_size_t), ("t_arrival_us", ctypes.c_double), ("t_submit_us", ctypes.c_double),
                ("t_launch_us", ctypes.c_double), ("t_done_us", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    _fields_ = [(n, _u64) for n in ("submitted", "admitted", "completed", "released", "rejected", "launches",
                                    "pending", "inflight_requests", "inflight_bytes", "peak_inflight_bytes",
                                    "budget_bytes", "cancelled", "expired")]


class SimConfig(ctypes.Structure):
    _fields_ = [("rate_per_s", ctypes.c_double), ("duration_s", ctypes.c_double), ("seed", _u64),
                ("closed_loop_depth", ctypes.c_int), ("pool", ctypes.POINTER(TextureDesc)),
                ("pool_count", ctypes.c_int), ("pool_expected", ctypes.POINTER(ctypes.c_void_p)),
                ("verify_every", ctypes.c_int), ("server", ServerConfig)]


class SimResult(ctypes.Structure):
    _fields_ = ([(n, _u64) for n in ("submitted", "completed", "rejected", "verified", "mismatches", "launches")] +
                [(n, ctypes.c_double) for n in ("wall_s", "megapixels", "mps", "offered_mps", "lat_mean_us",
                                                "lat_p50_us", "lat_p90_us", "lat_p99_us", "lat_max_us",
                                                "queue_p50_us", "queue_p99_us", "mean_textures_per_launch")] +
                [("peak_inflight_bytes", _u64)] +
                [(n, ctypes.c_double) for n in ("gen_lag_p50_us", "gen_lag_p99_us", "gen_lag_max_us",
                                                "submit_lat_p50_us", "submit_lat_p99_us")])

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


_lib.goofy_server_create.argtypes = [ctypes.POINTER(ServerConfig), ctypes.POINTER(_vp)]
_lib.goofy_server_submit.argtypes = [_vp, ctypes.POINTER(Request), ctypes.POINTER(_u64)]
_lib.goofy_server_poll.argtypes = [_vp, ctypes.POINTER(Completion), ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
_lib.goofy_server_release.argtypes = [_vp, _u64]
_lib.goofy_server_cancel.argtypes = [_vp, _u64]
_lib.goofy_server_wait_idle.argtypes = [_vp, ctypes.c_int]
_lib.goofy_server_stats_get.argtypes = [_vp, ctypes.POINTER(ServerStats)]
_lib.goofy_server_now_us.argtypes = [_vp]
_lib.goofy_server_now_us.restype = ctypes.c_double
_lib.goofy_server_destroy.argtypes = [_vp]
_lib.goofy_serve_simulate.argtypes = [ctypes.POINTER(SimConfig), ctypes.POINTER(SimResult)]


def server_config(budget: int = 32 << 30, arena: int = 0, streams: int = 4, max_batch: int = 256,
                  max_pending: int = 0, batching: bool = True, launch: dict | None = None) -> ServerConfig:
    c = ServerConfig(budget, arena, streams, max_batch, max_pending, 1 if batching else 0)
    if launch:
        c.launch = LaunchConfig(**launch)
    return c


class Server:
    """goofy_server: submit device-resident textures, poll completions, release outputs."""

    def __init__(self, config: ServerConfig | None = None, **kw):
        cfg = config if config is not None else server_config(**kw)
        h = _vp()
        _check(_lib.goofy_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h

    def submit(self, src, width: int, height: int, fmt: int, pitch: int | None = None, tag: int = 0,
               arrival_us: float = 0.0, deadline_us: float = 0.0) -> int:
        """Submit one request; arrival_us / deadline_us are on the server clock (now_us())."""
        r = Request(_ptr(src), width, height, width * 4 if pitch is None else pitch, fmt, tag, arrival_us,
                    deadline_us)
        t = _u64()
        _check(_lib.goofy_server_submit(self._h, ctypes.byref(r), ctypes.byref(t)))
        return int(t.value)

    def try_submit(self, src, width: int, height: int, fmt: int, pitch: int | None = None, tag: int = 0) -> int:
        """Like submit, but returns the status code instead of raising."""
        r = Request(_ptr(src), width, height, width * 4 if pitch is None else pitch, fmt, tag)
        t = _u64()
        return int(_lib.goofy_server_submit(self._h, ctypes.byref(r), ctypes.byref(t)))

    def cancel(self, ticket: int) -> int:
        """Cancel an unlaunched request (status code: 0, or -8 when already launched)."""
        return int(_lib.goofy_server_cancel(self._h, ticket))

    def now_us(self) -> float:
        return float(_lib.goofy_server_now_us(self._h))

    def poll(self, max_items: int = 256) -> list[Completion]:
        buf = (Completion * max_items)()
        n = ctypes.c_int()
        _check(_lib.goofy_server_poll(self._h, buf, max_items, ctypes.byref(n)))
        return [buf[i] for i in range(n.value)]

    def release(self, ticket: int) -> int:
        return int(_lib.goofy_server_release(self._h, ticket))

    def wait_idle(self, timeout_ms: int = -1) -> int:
        return int(_lib.goofy_server_wait_idle(self._h, timeout_ms))

    def stats(self) -> dict:
        s = ServerStats()
        _check(_lib.goofy_server_stats_get(self._h, ctypes.byref(s)))
        return {n: int(getattr(s, n)) for n, _ in s._fields_}

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.goofy_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def simulate(pool: list[dict], rate_per_s: float, duration_s: float, seed: int = 1, closed_loop_depth: int = 0,
             expected: list[np.ndarray] | None = None, verify_every: int = 0,
             server: ServerConfig | None = None) -> dict:
    """Run the load generator over device-resident pool textures; returns the result dict."""
    n = len(pool)
    arr = (TextureDesc * n)()
    for i, t in enumerate(pool):
        arr[i] = TextureDesc(_ptr(t["src"]), None, t["width"], t["height"], t.get("pitch", t["width"] * 4),
                             t["format"])
    exp_arr = None
    keep = []
    if expected is not None:
        exp_arr = (ctypes.c_void_p * n)()
        for i, e in enumerate(expected):
            e = np.ascontiguousarray(e, dtype=np.uint8)
            keep.append(e)
            exp_arr[i] = e.ctypes.data
    cfg = SimConfig(rate_per_s, duration_s, seed, closed_loop_depth, arr, n,
                    ctypes.cast(exp_arr, ctypes.POINTER(ctypes.c_void_p)) if exp_arr is not None else None,
                    verify_every, server if server is not None else server_config())
    res = SimResult()
    _check(_lib.goofy_serve_simulate(ctypes.byref(cfg), ctypes.byref(res)))
    return res.as_dict()


def make_pool(g, hip, textures: list[dict], stream=None):
    """Device-resident pool textures (one allocation) and their blocks encoded alone."""
    sizes = [(t["width"] * 4 * t["height"] + 255) // 256 * 256 for t in textures]
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    src = hip.DeviceBuffer(int(offs[-1]))
    osz = [(t["width"] // 4) * (t["height"] // 4) * g.block_bytes(t["format"]) for t in textures]
    ooffs = np.concatenate([[0], np.cumsum([(s + 255) // 256 * 256 for s in osz])]).astype(np.int64)
    alone = hip.DeviceBuffer(int(ooffs[-1]))
    pool = []
    for i, t in enumerate(textures):
        p = src.data_ptr() + int(offs[i])
        g.synth_device(t["kind"], t["seed"], t["texture_id"], t["noise"], p, t["width"], t["height"], stream=stream)
        g.compress_device(t["format"], alone.data_ptr() + int(ooffs[i]), p, t["width"], t["height"],
                          t["width"] * 4, stream)
        pool.append(dict(src=p, width=t["width"], height=t["height"], pitch=t["width"] * 4, format=t["format"]))
    hip.synchronize()
    expected = [alone.download(osz[i], int(ooffs[i])) for i in range(len(textures))]
    alone.free()
    return pool, expected, src


def bench_serving(g, hip, W, log=print, rate: float = 20000.0, duration_s: float = 10.0) -> dict:
    """The serving workload: fixed Poisson load with p50/p99, then capacity (closed loop)
    with batched launches vs one plain launch per request."""
    textures = W.serve_pool()
    log(f"serve: pool of {len(textures)} textures, generating on device")
    pool, expected, keep = make_pool(g, hip, textures)
    mean_mpx = float(np.mean([t["width"] * t["height"] for t in textures])) / 1e6
    log(f"serve: Poisson {rate:.0f} req/s for {duration_s:.0f} s")
    fixed = simulate(pool, rate, duration_s, seed=W.SERVE_SEED, expected=expected, verify_every=16,
                     server=server_config())
    log("serve: capacity, batched")
    cap_b = simulate(pool, 0.0, 2.0, seed=7, closed_loop_depth=512, server=server_config(batching=True))
    log("serve: capacity, one plain launch per request")
    cap_p = simulate(pool, 0.0, 2.0, seed=7, closed_loop_depth=512, server=server_config(batching=False))
    keep.free()
    r = lambda v, d=1: round(float(v), d)  # noqa: E731
    return {
        "workload": f"Poisson {rate:.0f} req/s, {duration_s:.0f} s, pool of {len(textures)} device-resident "
                    f"textures (W,H log-uniform multiples of 4 in [64,4096], mean {mean_mpx:.2f} Mpx), "
                    "50/50 DXT1/ETC1S, budget 32 GiB, 4 streams",
        "requests": int(fixed["completed"]), "offered_mps": r(fixed["offered_mps"]), "mps": r(fixed["mps"]),
        "latency_from": "scheduled Poisson arrival (includes the generator's lag)",
        "latency_us": {"p50": r(fixed["lat_p50_us"]), "p90": r(fixed["lat_p90_us"]), "p99": r(fixed["lat_p99_us"]),
                       "mean": r(fixed["lat_mean_us"]), "max": r(fixed["lat_max_us"])},
        "generator_lag_us": {"p50": r(fixed["gen_lag_p50_us"]), "p99": r(fixed["gen_lag_p99_us"]),
                             "max": r(fixed["gen_lag_max_us"])},
        "submit_to_done_us": {"p50": r(fixed["submit_lat_p50_us"]), "p99": r(fixed["submit_lat_p99_us"])},
        "queue_us": {"p50": r(fixed["queue_p50_us"]), "p99": r(fixed["queue_p99_us"])},
        "textures_per_launch": r(fixed["mean_textures_per_launch"], 2),
        "verified_requests": int(fixed["verified"]), "mismatches": int(fixed["mismatches"]),
        "peak_inflight_gib": r(fixed["peak_inflight_bytes"] / 2**30, 3),
        "capacity": {"batched_mps": r(cap_b["mps"]), "plain_per_request_mps": r(cap_p["mps"]),
                     "speedup": r(cap_b["mps"] / max(cap_p["mps"], 1e-9), 2),
                     "batched_p99_us": r(cap_b["lat_p99_us"]), "plain_p99_us": r(cap_p["lat_p99_us"]),
                     "textures_per_launch": r(cap_b["mean_textures_per_launch"], 2)},
    }
