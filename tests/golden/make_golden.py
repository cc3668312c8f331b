"""Writes tests/golden/ref_rays.npz: ray outputs of the reference's own code.

Inputs: seeded camera-sky directions for five poses (the C0 pose and four
others).  The initial conditions come from the oracle's restatement of
Grid::integration_wrapper (Grid.h:524-547, not compilable on its own); the
capture flag and the integration come from the reference's Metric.h
(checkCelest, Metric.h:169-174; rkckIntegrate1, Metric.h:307-324) compiled
from its sources into oracle/_ref/libbtref.so (oracle/Makefile, target ref).

Run in a checkout where oracle/_ref is built:  python tests/golden/make_golden.py
"""
import ctypes as C
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402

POSES = [(0.999, math.pi / 2, 0.0, 5.0), (0.5, 1.2, 0.3, 3.0), (0.9, 0.4, 1.0, 20.0),
         (0.3, 2.5, 5.0, 8.0), (0.001, math.pi / 2, 0.0, 10.0)]
N_RAYS = 1200


def main():
    R = O.load_ref()
    if R is None:
        sys.exit("oracle/_ref/libbtref.so is not built")
    out = {"poses": np.array(POSES)}
    rng = np.random.default_rng(20251202)
    for k, (a, th0, ph0, r) in enumerate(POSES):
        ps = O.pose(a, th0, ph0, r)
        t = np.arccos(1 - 2 * rng.random(N_RAYS))
        p = 2 * math.pi * rng.random(N_RAYS)
        ot = np.empty(N_RAYS)
        op = np.empty(N_RAYS)
        st = np.empty(N_RAYS, dtype=np.int32)
        for i in range(N_RAYS):
            pR, pT, b, q = O.initial_conditions(ps, t[i], p[i])
            if not R.btref_check_celest(a, pR, r, th0, b, q):
                ot[i], op[i], st[i] = -1.0, -1.0, 0
                continue
            to, po, so = C.c_double(), C.c_double(), C.c_int()
            R.btref_integrate(a, r, th0, ph0, pR, b, q, pT, C.byref(to), C.byref(po), C.byref(so))
            ot[i], op[i], st[i] = to.value, po.value, so.value
        out[f"p{k}_in_theta"] = t
        out[f"p{k}_in_phi"] = p
        out[f"p{k}_theta"] = ot
        out[f"p{k}_phi"] = op
        out[f"p{k}_step"] = st
        print(f"pose {k} {POSES[k]}: {np.sum(ot >= 0)} escaping of {N_RAYS}")
    np.savez_compressed(os.path.join(HERE, "ref_rays.npz"), **out)


if __name__ == "__main__":
    main()
