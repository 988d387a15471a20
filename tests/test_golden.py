"""The oracle against the committed golden records of one or two 3D steps
(tests/golden/*.npz, written by tests/golden/make_golden.py).

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


NAMES = sorted(make_golden.DECKS)


@pytest.mark.parametrize("name", NAMES)
def test_fixture_is_a_real_response(name):
    """The deck exercises the iteration (several slices need more than one
    iteration) and every mode of the fields."""
    golden = make_golden.load(name)
    assert golden["iters"].max() >= 4 and golden["iters"].min() >= 1
    assert len(golden["beam_x1"]) == 1 << 10
    b = golden["hist_b"]
    for k in range(b.shape[2]):
        assert np.abs(b[:, :, k, :]).max() > 1e-6 * np.abs(b).max(), f"mode array {k} is empty"


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_golden_step(name):
    golden = make_golden.load(name)
    rec = make_golden.oracle_record(name)
    assert sorted(rec) == sorted(golden)
    assert rec["iters"].tolist() == golden["iters"].tolist()
    ids = [k for k in golden if k.endswith("_id")]
    for k in ids:
        assert np.array_equal(rec[k], golden[k]), k
    bad = {k: rel_err(rec[k], golden[k]) for k in golden if k != "iters" and k not in ids}
    bad = {k: v for k, v in bad.items() if not v <= TOL}
    assert not bad, f"beyond {TOL}: {bad}"


def test_golden_check_sees_a_small_change():
    """The 1e-12 bar is tight enough to see a small change of the answer: a
    beam charge off by 1e-9 (relative) moves the record beyond it."""
    name = "step_small"
    golden = make_golden.load(name)
    d = make_golden.DECKS[name]()
    d["beam"][0]["total_charge"] *= 1.0 + 1e-9
    rec = make_golden.oracle_record(name, deck=d)
    worst = max(rel_err(rec[k], golden[k]) for k in ("hist_psi", "hist_b", "hist_q_beam"))
    assert worst > 100 * TOL, worst
