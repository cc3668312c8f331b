"""The oracle against the committed reference vectors (tests/golden/ref_rays.npz,
written by tests/golden/make_golden.py from the reference's own Metric.h).

Runs without the reference checkout: the fixture is data.  Bar: bit-exact
(capture flags, celestial angles and step counts)."""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_rays.npz"))
POSES = [tuple(p) for p in GOLD["poses"]]


@pytest.mark.parametrize("k", range(len(POSES)))
def test_oracle_reproduces_reference_vectors(k):
    a, th0, ph0, r = POSES[k]
    ps = O.pose(a, th0, ph0, r)
    t, p = GOLD[f"p{k}_in_theta"], GOLD[f"p{k}_in_phi"]
    ot, op, st, _ = O.trace_rays(ps, O.params(), t, p, nthreads=4)
    assert np.array_equal(ot, GOLD[f"p{k}_theta"])
    assert np.array_equal(op, GOLD[f"p{k}_phi"])
    assert np.array_equal(st, GOLD[f"p{k}_step"])
