"""The C-ABI library loads, exports every symbol include/bt_abi.h declares,
and its host-only entry points behave (no device compute here)."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

from blacktracer_amd import _lib
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(bt_[a-z0-9_]+)\s*\(", text):
            names.append(m.group(1))
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, f"declared but not exported: {missing}"
    bound = set(_lib.SIGNATURES)
    assert not (set(names) - bound), f"declared but not bound in _lib: {set(names) - bound}"


def test_version_and_defaults():
    L = _lib.load()
    assert L.bt_abi_version() == 2
    p = _lib.BtIntegParams()
    L.bt_integ_params_default(C.byref(p))
    assert (p.eps, p.h1, p.z_end, p.max_steps, p.apriori_capture) == (1e-5, 0.01, -1e7, 1000, 1)
    # the mixed path's routing defaults
    assert (p.crit_eps, p.hmin_fp64, p.noise_budget, p.decision_margin, p.drift_budget, p.polar_b) == \
        (0.03, 1e-6, 0.13, 4.0, 1.2e-6, 1e-3)
    g = _lib.BtGridParams()
    L.bt_grid_params_default(C.byref(g))
    assert g.refine_thresh == 0.015 and g.fix_tvertices == 1


@pytest.mark.parametrize("pose", [(0.999, math.pi / 2, 0.0, 5.0), (0.5, 1.2, 0.3, 3.0),
                                  (0.9, 0.4, 1.0, 20.0)])
def test_camera_init_matches_oracle_bit_exact(pose):
    a, th, ph, r = pose
    L = _lib.load()
    bp = _lib.BtPose(a=a, theta=th, phi=ph, r=r, br=0.0, btheta=0.0, bphi=1.0)
    out = (C.c_double * 6)()
    cam7 = (C.c_float * 7)()
    assert L.bt_camera_init(C.byref(bp), out, cam7) == 0
    ps = O.pose(a, th, ph, r)
    assert list(out) == [ps.speed, ps.alpha, ps.w, ps.wbar, ps.Delta, ps.ro]
    assert np.array_equal(np.array(list(cam7), dtype=np.float32),
                          np.array([ps.speed, ps.alpha, ps.w, ps.wbar, 0.0, 0.0, 1.0],
                                   dtype=np.float32))


def test_invalid_arguments_are_rejected():
    L = _lib.load()
    bp = _lib.BtPose(a=1.5, theta=1.0, phi=0.0, r=5.0, bphi=1.0)
    assert L.bt_camera_init(C.byref(bp), None, None) == _lib.BT_ERR_INVALID
    assert b"|a|" in L.bt_last_error()
    bp = _lib.BtPose(a=0.5, theta=1.0, phi=0.0, r=1.5, bphi=1.0)   # inside the horizon
    assert L.bt_camera_init(C.byref(bp), None, None) == _lib.BT_ERR_INVALID
    assert L.bt_context_create(0, None) == _lib.BT_ERR_INVALID
    assert L.bt_grid_info(None, None, None, None, None, None, None) == _lib.BT_ERR_INVALID


def test_context_without_device_fails_cleanly():
    """On a machine without a HIP device the context reports BT_ERR_DEVICE."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    L = _lib.load()
    h = C.c_void_p()
    assert L.bt_context_create(0, C.byref(h)) == _lib.BT_ERR_DEVICE
    assert not h.value
