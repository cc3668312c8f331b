"""Launch-shape tuning (bt_tune_set / bt_autotune): the traced angles and
step indices are the same bit for bit under every launch shape, FP64 and
mixed, since each ray's arithmetic does not depend on which lane runs it."""
import math

import numpy as np
import pytest

import blacktracer_amd as bt

pytestmark = pytest.mark.gpu

SHAPES = [  # (chunk, block, waves)
    (128, 256, 0),
    (64, 64, 1),
    (1024, 128, 2),
    (256, 256, 4),
    (512, 64, 0),
]


def _rays(n, seed):
    rng = np.random.default_rng(seed)
    th = np.arccos(1.0 - 2.0 * rng.random(n))
    ph = 2.0 * math.pi * rng.random(n)
    return th, ph


@pytest.mark.parametrize("precision", [bt.BT_PREC_FP64, bt.BT_PREC_MIXED])
def test_results_independent_of_launch_shape(precision):
    ctx = bt.Context(0)
    try:
        bh = bt.BlackHole(0.999)
        cam = bt.Camera(math.pi / 2, 0.0, 5.0, black_hole=bh)
        th, ph = _rays(40000, 7)
        prm = bt.IntegParams(precision)
        ref = None
        for shape in SHAPES:
            print(f"precision={precision} shape={shape}", flush=True)
            ctx.set_tune(*shape)
            assert ctx.tune() == dict(zip(("chunk", "block", "waves"), shape))
            out = bt.trace_rays(cam, th, ph, params=prm, ctx=ctx)
            if ref is None:
                ref = out
                continue
            for a, b in zip(out, ref):
                assert np.array_equal(a, b)
    finally:
        ctx.close()


def test_tune_rejects_bad_shapes():
    ctx = bt.Context(0)
    try:
        for bad in [(128, 96, 0), (32, 256, 0), (128, 256, -1), (128, 256, 9)]:
            with pytest.raises(bt.BtError):
                ctx.set_tune(*bad)
        assert ctx.tune() == {"chunk": 128, "block": 256, "waves": 0}
    finally:
        ctx.close()


def test_autotune_keeps_a_valid_fastest_shape():
    ctx = bt.Context(0)
    try:
        bh = bt.BlackHole(0.999)
        cam = bt.Camera(math.pi / 2, 0.0, 5.0, black_hole=bh)
        prm = bt.IntegParams(bt.BT_PREC_MIXED)
        th, ph = _rays(20000, 3)
        before = bt.trace_rays(cam, th, ph, params=prm, ctx=ctx)
        res = ctx.autotune(cam, params=prm, n_rays=1 << 17)
        print(f"autotune {res}", flush=True)
        assert res["block"] in (64, 128, 256)
        assert 64 <= res["chunk"] <= 1024
        assert 0 <= res["waves"] <= 4
        assert res["ms"] > 0 and res["rays_per_lane"] > 0
        assert ctx.tune() == {k: res[k] for k in ("chunk", "block", "waves")}
        after = bt.trace_rays(cam, th, ph, params=prm, ctx=ctx)
        for a, b in zip(after, before):
            assert np.array_equal(a, b)
    finally:
        ctx.close()
