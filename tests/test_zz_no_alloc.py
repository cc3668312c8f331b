"""bt_trace_rays_device_noalloc (the FP64 ray batch with its work queue in
the library's fixed device array) against bt_trace_rays_device, bit for bit,
through the C ABI: two calls in a row, the second smaller, so that the
second must not see the first call's queue or counters."""
import ctypes as C
import math

import numpy as np
import pytest

from blacktracer_amd import _lib

pytestmark = pytest.mark.gpu

L = None


def lib():
    global L
    if L is None:
        L = _lib.load()
        L.bt_trace_rays_device_noalloc.restype = C.c_int
        L.bt_trace_rays_device_noalloc.argtypes = [C.c_void_p, C.POINTER(_lib.BtPose),
                                                   C.POINTER(_lib.BtIntegParams), C.c_int64,
                                                   C.c_void_p, C.c_void_p, C.c_void_p]
    return L


def ok(rc):
    assert rc == _lib.BT_OK, lib().bt_last_error().decode()


class Dev:
    def __init__(self, ctx, arr):
        self.ctx, self.nbytes = ctx, arr.nbytes
        self.ptr = C.c_void_p()
        ok(lib().bt_device_alloc(ctx, self.nbytes, C.byref(self.ptr)))
        ok(lib().bt_memcpy(ctx, self.ptr, arr.ctypes.data_as(C.c_void_p), self.nbytes, 1))

    def get(self, like):
        out = np.empty_like(like)
        ok(lib().bt_memcpy(self.ctx, out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes, 0))
        return out

    def free(self):
        lib().bt_device_free(self.ctx, self.ptr)


def directions(n, seed):
    rng = np.random.default_rng(seed)
    return np.arccos(1 - 2 * rng.random(n)), 2 * math.pi * rng.random(n)


def run(ctx, entry, pose, prm, th, ph):
    st = np.zeros(th.size, np.int32)
    dt, dp, ds = Dev(ctx, th), Dev(ctx, ph), Dev(ctx, st)
    if entry == "noalloc":
        rc = lib().bt_trace_rays_device_noalloc(ctx, C.byref(pose), C.byref(prm), th.size, dt.ptr, dp.ptr, ds.ptr)
    else:
        rc = lib().bt_trace_rays_device(ctx, C.byref(pose), C.byref(prm), th.size, dt.ptr, dp.ptr, ds.ptr, None)
    ok(rc)
    ok(lib().bt_context_sync(ctx))
    out = dt.get(th), dp.get(ph), ds.get(st)
    for d in (dt, dp, ds):
        d.free()
    return out


@pytest.fixture(scope="module")
def ctx():
    h = C.c_void_p()
    ok(lib().bt_context_create(0, C.byref(h)))
    yield h
    lib().bt_context_destroy(h)


def test_noalloc_equals_device_entry_twice_second_smaller(ctx):
    pose = _lib.BtPose(a=0.999, theta=math.pi / 2, phi=0.0, r=5.0, br=0.0, btheta=0.0, bphi=1.0)
    prm = _lib.BtIntegParams()
    lib().bt_integ_params_default(C.byref(prm))
    for n, seed in ((200_000, 1), (7_001, 2)):
        th, ph = directions(n, seed)
        print(f"noalloc vs device entry: n={n} seed={seed}", flush=True)
        a = run(ctx, "device", pose, prm, th, ph)
        b = run(ctx, "noalloc", pose, prm, th, ph)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
        assert (a[0] == -1).any() and (a[0] != -1).any()   # both captured and escaping rays


def test_noalloc_rejects_oversize_and_mixed(ctx):
    pose = _lib.BtPose(a=0.9, theta=1.0, phi=0.0, r=10.0, br=0.0, btheta=0.0, bphi=1.0)
    prm = _lib.BtIntegParams()
    lib().bt_integ_params_default(C.byref(prm))
    dummy = C.c_void_p(16)   # never dereferenced: both calls fail their checks first
    rc = lib().bt_trace_rays_device_noalloc(ctx, C.byref(pose), C.byref(prm), (1 << 20) + 1, dummy, dummy, dummy)
    assert rc == _lib.BT_ERR_INVALID and b"exceeds" in lib().bt_last_error()
    prm.precision = 1
    rc = lib().bt_trace_rays_device_noalloc(ctx, C.byref(pose), C.byref(prm), 10, dummy, dummy, dummy)
    assert rc == _lib.BT_ERR_INVALID
