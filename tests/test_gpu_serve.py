"""Continuous-batching render server (bt_server_*).

Bar: a request's frame is byte-identical alone, in a mix of concurrent
requests of other poses and resolutions, and to bt_render_frame on a grid
built with the same settings; admission never exceeds the slots or (when a
frame and a grid fit) the HBM budget; a frame larger than the budget is
rejected at submit."""
import math

import numpy as np
import pytest

import blacktracer_amd as bt
from blacktracer_amd import scene as S

pytestmark = pytest.mark.gpu

LEVEL = 7
POSES = [(0.999, 5.0, math.pi / 2), (0.6, 8.0, 1.2), (0.9, 15.0, 2.0)]


@pytest.fixture(scope="module")
def scene(ctx):
    stars = S.synthetic_stars(20000, seed=3)
    sky = S.synthetic_sky(512, seed=4)
    return bt.Scene(stars, sky, tree_level=8, ctx=ctx)


def _run(server, reqs, timeout=120.0):
    outs = {}
    for tag, (pose, off, (w, h)) in enumerate(reqs):
        out = np.zeros((h, w, 4), dtype=np.uint8)
        outs[tag] = out
        server.submit(*pose, off, out, tag)
    server.drain(timeout)
    comps = server.poll(1024)
    assert len(comps) == len(reqs)
    assert all(c["status"] == 0 for c in comps)
    return outs, comps


def test_alone_equals_in_mix_equals_direct(ctx, scene):
    x = (POSES[0], 0.37, (320, 180))
    srv = bt.Server(scene, slots=4, max_level=LEVEL, ctx=ctx)
    alone, _ = _run(srv, [x])
    mix = []
    for k in range(12):
        mix.append((POSES[k % 3], 0.05 * k, [(160, 96), (320, 180), (640, 360)][k % 3]))
    mix.insert(5, x)
    mixed, comps = _run(srv, mix)
    st = srv.stats()
    print(f"mix: peak in flight {st['peak_inflight']}, builds {st['grid_builds']}, hits {st['grid_hits']}",
          flush=True)
    assert np.array_equal(alone[0], mixed[5])
    assert st["peak_inflight"] >= 2          # frames did overlap
    srv.close()
    # the same frame through the resident render entry point
    bh = bt.BlackHole(x[0][0])
    cam = bt.Camera(x[0][2], 0.0, x[0][1], black_hole=bh)
    g = bt.Grid(LEVEL, 1, True, cam, bh, ctx=ctx)
    img = bt.render_frame(g, scene, bt.View(320, 180, ctx=ctx), flags=bt.RenderFlags(offset=x[1]))
    assert np.array_equal(img, alone[0])


def test_budget_admission_and_eviction(ctx, scene):
    # room for one 640x360 slot's buffers, its view, a grid build and about
    # one and a half level-7 grids: frames run one at a time and the three
    # poses evict each other
    fp = bt.server_footprint(640, 360, max_level=LEVEL)
    budget = fp["slot"] + fp["view"] + fp["build"] + 3 * fp["grid"] // 2
    srv = bt.Server(scene, hbm_budget=budget, slots=4, max_level=LEVEL, ctx=ctx)
    reqs = [(POSES[k % 3], 0.1 * k, (640, 360)) for k in range(9)]
    outs, comps = _run(srv, reqs)
    st = srv.stats()
    print(f"budget: peak bytes {st['peak_bytes'] >> 20} MiB, in flight {st['peak_inflight']}, "
          f"evictions {st['evictions']}", flush=True)
    assert st["completed"] == 9
    assert st["peak_bytes"] <= budget
    assert st["peak_inflight"] == 1
    assert st["evictions"] > 0
    srv.close()
    # every frame equals the same request served alone
    ref = bt.Server(scene, slots=1, max_level=LEVEL, ctx=ctx)
    for k in (0, 4, 8):
        alone, _ = _run(ref, [reqs[k]])
        assert np.array_equal(alone[0], outs[k])
    ref.close()
    with pytest.raises(bt.BtError):
        tiny = bt.Server(scene, hbm_budget=1 << 20, slots=1, max_level=LEVEL, ctx=ctx)
        tiny.submit(*POSES[0], 0.0, np.zeros((360, 640, 4), dtype=np.uint8), 0)
    tiny.close()


def test_request_that_can_never_fit_fails(ctx, scene):
    """Frame + view + grid above the budget with nothing in flight: the
    request completes with BT_ERR_BUDGET instead of waiting forever."""
    # the frame alone fits; its slot's buffers, its view and its grid do not
    fp = bt.server_footprint(640, 360, max_level=LEVEL)
    srv = bt.Server(scene, hbm_budget=fp["slot"] + fp["view"] // 2, slots=2, max_level=LEVEL, ctx=ctx)
    out = np.zeros((360, 640, 4), dtype=np.uint8)
    srv.submit(*POSES[0], 0.0, out, 7)
    srv.drain(60)
    comps = srv.poll(8)
    assert len(comps) == 1 and comps[0]["status"] == bt.BT_ERR_BUDGET
    srv.close()


def _hip_free_bytes():
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def test_device_memory_stays_within_the_budget(ctx, scene):
    """The budget counts what the device really holds: the drop in free
    device memory (hipMemGetInfo) over a budget-limited run, sampled while
    frames run and after they drained, stays within hbm_budget.  The budget
    holds two 640 x 360 slots' work buffers, not four, so slots are trimmed
    and grids evicted."""
    import time
    fp = bt.server_footprint(640, 360, max_level=LEVEL)
    budget = 2 * fp["slot"] + 2 * fp["view"] + fp["build"] + 2 * fp["grid"]
    srv = bt.Server(scene, hbm_budget=budget, slots=4, max_level=LEVEL, ctx=ctx)
    ctx.sync()
    time.sleep(0.05)
    free0 = _hip_free_bytes()
    poses = POSES + [(0.3, 6.0, 1.0)]
    reqs = [(poses[k % 4], 0.07 * k, [(640, 360), (320, 180)][k % 2]) for k in range(16)]
    outs = []
    for tag, (pose, off, (w, h)) in enumerate(reqs):
        out = np.zeros((h, w, 4), dtype=np.uint8)
        outs.append(out)
        srv.submit(*pose, off, out, tag)
    low = free0
    t0 = time.perf_counter()
    done = 0
    while done < len(reqs) and time.perf_counter() - t0 < 120:
        low = min(low, _hip_free_bytes())
        done += len(srv.poll(64))
        time.sleep(0.002)
    srv.drain(60)
    low = min(low, _hip_free_bytes())
    st = srv.stats()
    drop = free0 - low
    print(f"budget {budget >> 20} MiB: free-memory drop {drop >> 20} MiB, counted peak {st['peak_bytes'] >> 20} MiB, "
          f"evictions {st['evictions']}, completed {st['completed']}", flush=True)
    assert st["completed"] == len(reqs)
    assert st["peak_bytes"] <= budget
    assert drop <= budget
    srv.close()
