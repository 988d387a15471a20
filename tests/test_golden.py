"""The oracle against the committed golden record of one 3D step
(tests/golden/step_small.npz, written by tests/golden/make_golden.py).

The record pins the oracle's whole-step answer on a fixed deck: plasma and
beam injection, every slice's fields and iteration count, the plasma push and
the beam push.  A change to the oracle that moves any of them fails here,
before it can move the reference the GPU is checked against."""
import os
import sys

import numpy as np
import pytest

from conftest import rel_err

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden  # noqa: E402

# the oracle is serial FP64 C built with the same flags: any difference beyond
# the last few bits is a change of the algorithm
TOL = 1e-12


@pytest.fixture(scope="module")
def golden():
    with np.load(make_golden.FIXTURE) as z:
        return {k: z[k] for k in z.files}


def test_fixture_is_a_real_response(golden):
    """The deck exercises the iteration (several slices need more than one
    iteration) and every mode of the fields."""
    assert golden["iters"].max() >= 4 and golden["iters"].min() >= 1
    assert len(golden["beam_x1"]) == 1 << 10
    b = golden["hist_b"]
    for k in range(b.shape[2]):
        assert np.abs(b[:, :, k, :]).max() > 1e-6 * np.abs(b).max(), f"mode array {k} is empty"


def test_oracle_reproduces_golden_step(golden):
    rec = make_golden.oracle_record()
    assert sorted(rec) == sorted(golden)
    assert rec["iters"].tolist() == golden["iters"].tolist()
    assert np.array_equal(rec["sp_id"], golden["sp_id"])
    bad = {k: rel_err(rec[k], golden[k]) for k in golden if k not in ("iters", "sp_id")}
    bad = {k: v for k, v in bad.items() if not v <= TOL}
    assert not bad, f"beyond {TOL}: {bad}"
