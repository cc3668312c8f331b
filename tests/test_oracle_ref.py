"""Pins the CPU oracle (Path A) against the reference's own code.

oracle/_ref/libbtref.so is the reference's Metric.h / Camera.h / BlackHole.h
compiled from the reference checkout (oracle/Makefile, target `ref`).  The
oracle is a separate plain-C restatement; both are built with
-ffp-contract=off against the same libm, so every result must agree
bit-for-bit.  Skipped where the reference checkout (and so _ref) is absent.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import oracle as O

R = O.load_ref()
pytestmark = pytest.mark.skipif(R is None, reason="oracle/_ref not built (no reference checkout)")

POSES = [(0.999, math.pi / 2, 0.0, 5.0), (0.5, 1.2, 0.3, 3.0), (0.9, 0.4, 1.0, 20.0),
         (0.3, 2.5, 5.0, 8.0), (0.001, math.pi / 2, 0.0, 10.0)]


def _dirs(n, seed):
    rng = np.random.default_rng(seed)
    return np.arccos(1 - 2 * rng.random(n)), 2 * math.pi * rng.random(n)


@pytest.mark.parametrize("pose", POSES)
def test_pose_bit_exact(pose):
    a, th, ph, r = pose
    ps = O.pose(a, th, ph, r)
    out = np.zeros(6)
    R.btref_pose(a, th, ph, r, 0.0, 0.0, 1.0, O._ptr(out, O.DP))
    assert [ps.speed, ps.alpha, ps.w, ps.wbar, ps.Delta, ps.ro] == list(out)
    assert O.lib().bto_calc_speed(a, r, th) == R.btref_calc_speed(a, r, th)


def test_b0_q0_bit_exact():
    L = O.lib()
    for a in (0.001, 0.3, 0.9, 0.999):
        for r0 in np.linspace(1.05, 6.0, 57):
            assert L.bto_b0(a, r0) == R.btref_b0(a, r0)
            assert L.bto_q0(a, r0) == R.btref_q0(a, r0)


@pytest.mark.parametrize("pose", POSES[:4])
def test_capture_test_bit_exact(pose):
    a, th, ph, r = pose
    ps = O.pose(a, th, ph, r)
    L = O.lib()
    t, p = _dirs(3000, 11)
    for i in range(t.size):
        pR, pT, b, q = O.initial_conditions(ps, t[i], p[i])
        nw, gd = C.c_int64(), C.c_int64()
        mine = L.bto_check_celest(a, pR, r, th, b, q, 100000, C.byref(nw), C.byref(gd))
        assert mine == R.btref_check_celest(a, pR, r, th, b, q)
        assert L.bto_check_bq(a, b, q, 100000, None) == R.btref_check_bq(a, b, q)
        assert L.bto_check_rup(a, r, th, b, q, None) == R.btref_check_rup(a, r, th, b, q)


def test_derivs_and_rkck_bit_exact():
    L = O.lib()
    rng = np.random.default_rng(3)
    for _ in range(500):
        a = rng.uniform(0.001, 0.999)
        var = np.array([rng.uniform(1.5, 50), rng.uniform(0.05, 3.1), rng.uniform(0, 6.28),
                        rng.uniform(-2, 2), rng.uniform(-5, 5)])
        b, q = rng.uniform(-8, 8), rng.uniform(0, 40)
        o1, o2 = np.zeros(5), np.zeros(5)
        L.bto_derivs(a, O._ptr(var, O.DP), O._ptr(o1, O.DP), b, q)
        R.btref_derivs(a, O._ptr(var, O.DP), O._ptr(o2, O.DP), b, q)
        assert np.array_equal(o1, o2)
        h = -10 ** rng.uniform(-3, 0)
        y1, e1, y2, e2 = np.zeros(5), np.zeros(5), np.zeros(5), np.zeros(5)
        L.bto_rkck(a, O._ptr(var, O.DP), O._ptr(o1, O.DP), h, O._ptr(y1, O.DP), O._ptr(e1, O.DP), b, q)
        R.btref_rkck(a, O._ptr(var, O.DP), O._ptr(o1, O.DP), h, O._ptr(y2, O.DP), O._ptr(e2, O.DP), b, q)
        assert np.array_equal(y1, y2) and np.array_equal(e1, e2)


@pytest.mark.parametrize("pose", POSES)
def test_rays_bit_exact(pose):
    """integration_wrapper's loop body: oracle vs reference rkckIntegrate1."""
    a, th, ph, r = pose
    ps = O.pose(a, th, ph, r)
    t, p = _dirs(1500, 5)
    ot, op, ost, _ = O.trace_rays(ps, O.params(), t, p, nthreads=4)
    n_int = 0
    for i in range(t.size):
        pR, pT, b, q = O.initial_conditions(ps, t[i], p[i])
        if not R.btref_check_celest(a, pR, r, th, b, q):
            assert ot[i] == -1 and op[i] == -1 and ost[i] == 0
            continue
        n_int += 1
        to, po, so = C.c_double(), C.c_double(), C.c_int()
        R.btref_integrate(a, r, th, ph, pR, b, q, pT, C.byref(to), C.byref(po), C.byref(so))
        assert (to.value, po.value, so.value) == (ot[i], op[i], ost[i])
    assert n_int > 100


def test_reference_batch_equals_oracle():
    """The reference's batch loop (threads over rays, the CPU baseline of the
    benchmark) gives the oracle's results bit for bit."""
    a, th, ph, r = POSES[0]
    ps = O.pose(a, th, ph, r)
    t, p = _dirs(800, 11)
    ot, op, ost, _ = O.trace_rays(ps, O.params(), t, p, nthreads=4)
    ic = O.initial_conditions_batch(ps, t, p)
    rt, rp, rs = O.ref_trace_batch(a, r, th, ph, ic, nthreads=4)
    assert np.array_equal(rt, ot) and np.array_equal(rp, op) and np.array_equal(rs, ost)


def test_wrap_to_pi_bit_exact():
    L = O.lib()
    rng = np.random.default_rng(9)
    for x, y in zip(rng.uniform(-40, 40, 2000), rng.uniform(-80, 80, 2000)):
        a1, b1 = C.c_double(x), C.c_double(y)
        a2, b2 = C.c_double(x), C.c_double(y)
        L.bto_wrap_to_pi(C.byref(a1), C.byref(b1))
        R.btref_wrap_to_pi(C.byref(a2), C.byref(b2))
        assert (a1.value, b1.value) == (a2.value, b2.value)
