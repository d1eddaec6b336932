"""Serving: continuous batching under an HBM budget (GPU, through the C-ABI).

Every result must be bit-identical to the same texture encoded alone, both
at the full request rate and under a budget tight enough to force queueing.
"""
import sys
import time

import numpy as np
import pytest

from goofy_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def pool(gpu):
    from goofy_amd import hip, serving
    ts = W.serve_pool(n=48, lo=4, hi=640, seed=99)
    say("serving pool: 48 textures")
    pool, expected, keep = serving.make_pool(gpu, hip, ts)
    yield ts, pool, expected
    keep.free()


def test_pool_alone_matches_oracle(gpu, oracle, pool):
    ts, p, expected = pool
    for i in range(0, len(ts), 5):
        t = ts[i]
        img = gpu.synth_host(t["kind"], t["seed"], t["texture_id"], t["noise"], t["width"], t["height"])
        assert np.array_equal(expected[i], oracle.compress_blocks(t["format"], img, t["width"], t["height"],
                                                                  t["width"] * 4)), t


def test_alone_equals_mix_at_full_rate(gpu, pool):
    from goofy_amd import serving
    ts, p, expected = pool
    say("serving: Poisson 20k req/s, 0.5 s, every result verified")
    res = serving.simulate(p, 20000.0, 0.5, seed=3, expected=expected, verify_every=1)
    say(res)
    assert res["submitted"] > 5000 and res["completed"] == res["submitted"]
    assert res["verified"] == res["completed"] and res["mismatches"] == 0
    assert res["lat_p99_us"] > 0 and res["lat_p50_us"] <= res["lat_p99_us"]


def test_alone_equals_mix_plain_mode_and_tight_budget(gpu, pool):
    from goofy_amd import serving
    ts, p, expected = pool
    biggest = max(t["width"] * t["height"] * 4 + (t["width"] // 4) * (t["height"] // 4) * 8 + 256 for t in ts)
    for batching in (True, False):
        cfg = serving.server_config(budget=biggest * 2, arena=biggest, streams=2, max_batch=8, batching=batching)
        say(f"serving: closed loop, budget {biggest * 2} B, batching={batching}")
        res = serving.simulate(p, 0.0, 0.3, seed=5, closed_loop_depth=64, expected=expected, verify_every=1,
                               server=cfg)
        assert res["completed"] == res["submitted"] > 50
        assert res["mismatches"] == 0 and res["verified"] == res["completed"]
        assert res["peak_inflight_bytes"] <= biggest * 2


@pytest.mark.parametrize("launch", [dict(blocks_per_lane=2, wg_waves=4, waves_per_cu=4, nontemporal=1, tile_order=1),
                                    dict(blocks_per_lane=4, wg_waves=1, waves_per_cu=0, nontemporal=0, tile_order=0)])
def test_alone_equals_mix_with_explicit_launch_shapes(gpu, pool, launch):
    """A server configured with a fixed shape (persistent grid and XCD order,
    or 4 blocks per lane) instead of the per-launch default."""
    from goofy_amd import serving
    ts, p, expected = pool
    say(f"serving: closed loop with launch {launch}")
    res = serving.simulate(p, 0.0, 0.3, seed=11, closed_loop_depth=96, expected=expected, verify_every=1,
                           server=serving.server_config(launch=launch))
    assert res["completed"] == res["submitted"] > 50
    assert res["mismatches"] == 0 and res["verified"] == res["completed"]


def test_budget_admission_and_backpressure(gpu, pool):
    from goofy_amd import hip, serving
    import goofy_amd as g
    ts, p, expected = pool
    i = max(range(len(ts)), key=lambda k: ts[k]["width"] * ts[k]["height"])
    t, src = ts[i], p[i]["src"]
    need = t["width"] * t["height"] * 4 + ((t["width"] // 4) * (t["height"] // 4) * 8 + 255) // 256 * 256
    # room for exactly one request in flight; max 3 waiting
    srv = serving.Server(serving.server_config(budget=need, arena=need, streams=2, max_pending=3))
    # nothing is released yet, so at most 1 admitted + 3 pending can be accepted
    tickets, busy = [], 0
    for k in range(6):
        if srv.try_submit(src, t["width"], t["height"], t["format"], tag=k) == 0:
            tickets.append(k)
        else:
            busy += 1
    assert busy >= 2 and 3 <= len(tickets) <= 4
    done = []
    deadline = time.time() + 20
    while len(done) < len(tickets) and time.time() < deadline:
        c = srv.poll()
        for x in c:
            assert x.status == 0
            hip.synchronize()
            got = np.empty(x.dst_bytes, np.uint8)
            hip.check(hip._hip.hipMemcpy(got.ctypes.data, x.dst, x.dst_bytes, hip.D2H), "D2H")
            assert np.array_equal(got, expected[i])
            done.append(x.ticket)
            st = srv.stats()
            assert st["inflight_bytes"] <= need
            # the next request is only admitted once this output is released
            assert srv.release(x.ticket) == 0
            assert srv.release(x.ticket) == g.ERR_INVALID  # double release
        time.sleep(0.001)
    assert len(done) == len(tickets)
    st = srv.stats()
    n = len(tickets)
    assert st["completed"] == n and st["released"] == n and st["rejected"] == busy
    assert st["peak_inflight_bytes"] <= need
    srv.close()


def test_invalid_requests(gpu, pool):
    from goofy_amd import serving
    import goofy_amd as g
    ts, p, expected = pool
    srv = serving.Server(serving.server_config(budget=1 << 20, arena=1 << 19))
    src = p[0]["src"]
    assert srv.try_submit(src, 6, 4, 0) == g.ERR_WIDTH
    assert srv.try_submit(src, 8, 6, 0) == g.ERR_HEIGHT
    assert srv.try_submit(src, 8, 4, 5) == g.ERR_INVALID
    assert srv.try_submit(src + 4, 8, 4, 0) == g.ERR_INVALID       # misaligned
    assert srv.try_submit(src, 1024, 1024, 0) == g.ERR_INVALID     # larger than the budget
    assert srv.release(12345) == g.ERR_INVALID
    assert srv.wait_idle(1000) == 0
    srv.close()


def test_cancel_and_deadlines_on_the_gpu(gpu, pool):
    """One stream, one texture per launch: a request cancelled while it waits
    completes with GOOFY_ERR_CANCELLED and holds no output; one whose
    deadline has passed completes with GOOFY_ERR_DEADLINE unlaunched; the
    rest encode exactly as alone."""
    from goofy_amd import hip, serving
    import goofy_amd as g
    ts, p, expected = pool
    i = max(range(len(ts)), key=lambda k: ts[k]["width"] * ts[k]["height"])
    t = ts[i]
    srv = serving.Server(serving.server_config(streams=1, max_batch=1))
    now = srv.now_us()
    say(f"serving: cancel and deadlines, texture {t['width']}x{t['height']}")
    tickets = [srv.submit(p[i]["src"], t["width"], t["height"], t["format"], tag=k,
                          deadline_us=(now - 1.0 if k == 2 else (now + 60e6 if k == 3 else 0.0))) for k in range(6)]
    rc = srv.cancel(tickets[5])
    assert rc in (0, g.ERR_BUSY)  # cancelled while waiting, or (unlikely) already launched
    assert srv.cancel(10 ** 9) == g.ERR_INVALID
    got = {}
    deadline = time.time() + 20
    while len(got) < 6 and time.time() < deadline:
        for c in srv.poll():
            got[c.ticket] = (c.status, c.dst, c.dst_bytes, c.t_launch_us)
            if c.status == 0:
                hip.synchronize()
                out = np.empty(c.dst_bytes, np.uint8)
                hip.check(hip._hip.hipMemcpy(out.ctypes.data, c.dst, c.dst_bytes, hip.D2H), "D2H")
                assert np.array_equal(out, expected[i])
                assert srv.release(c.ticket) == 0
        time.sleep(0.001)
    assert len(got) == 6
    assert got[tickets[2]][0] == g.ERR_DEADLINE and not got[tickets[2]][1] and got[tickets[2]][3] == 0
    assert got[tickets[3]][0] == 0
    assert got[tickets[5]][0] == (g.ERR_CANCELLED if rc == 0 else 0)
    for k in (0, 1, 4):
        assert got[tickets[k]][0] == 0
    st = srv.stats()
    assert st["expired"] == 1 and st["cancelled"] == (1 if rc == 0 else 0) and st["inflight_bytes"] == 0
    assert srv.wait_idle(10000) == 0
    srv.close()


def test_serving_workload_at_full_size(gpu, oracle):
    """The serving workload at its stated size: the 512-texture pool
    (independent W, H log-uniform over multiples of 4 in [64, 4096], 50/50
    DXT1 / ETC1S), Poisson 20k requests/s for 1 s under the 32 GiB budget,
    every result copied back and compared with the texture encoded alone; and
    the pool's encoding against the oracle on 32 textures, the 8 largest
    among them."""
    from goofy_amd import hip, serving
    ts = W.serve_pool()
    say(f"serving at full size: generating and encoding the pool of {len(ts)} textures alone")
    p, expected, keep = serving.make_pool(gpu, hip, ts)
    try:
        say("serving at full size: Poisson 20k req/s for 1 s, every result verified")
        res = serving.simulate(p, 20000.0, 1.0, seed=W.SERVE_SEED, expected=expected, verify_every=1,
                               server=serving.server_config(budget=32 << 30))
        say({k: round(v, 1) if isinstance(v, float) else v for k, v in res.items()})
        assert res["submitted"] > 15000 and res["rejected"] == 0
        assert res["completed"] == res["submitted"] and res["verified"] == res["completed"]
        assert res["mismatches"] == 0
        assert 0 < res["lat_p50_us"] <= res["lat_p99_us"] and res["peak_inflight_bytes"] <= 32 << 30
        assert res["gen_lag_p50_us"] >= 0
    finally:
        keep.free()
    areas = [t["width"] * t["height"] for t in ts]
    largest = sorted(range(len(ts)), key=lambda k: -areas[k])[:8]
    others = [k for k in range(0, len(ts), 21) if k not in largest][:24]
    say(f"serving at full size: {len(largest) + len(others)} pool textures against the oracle, "
        f"largest {ts[largest[0]]['width']}x{ts[largest[0]]['height']}")
    for k in largest + others:
        t = ts[k]
        img = gpu.synth_host(t["kind"], t["seed"], t["texture_id"], t["noise"], t["width"], t["height"])
        want = oracle.compress_blocks(t["format"], img, t["width"], t["height"], t["width"] * 4, threads=8)
        assert np.array_equal(expected[k], want), t
