"""Write the golden record of one 3D step of a small deck, from the oracle.

QPAD ships no golden vectors for this path (its test programs hold no
assertions), so the fixture is the oracle's own output on a fixed deck,
committed after the oracle was pinned against mathematics.  It serves two
checks: the oracle must keep reproducing it (a change to the oracle that moves
its answer fails a CPU test), and the HIP library must match it on a GPU.

    python tests/golden/make_golden.py      # rewrites step_small.npz
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]

FIXTURE = os.path.join(HERE, "step_small.npz")


def deck():
    """Nr 32, 24 slices, modes 0..2, 8 particles per cell, a 1024-particle
    beam (sigma_xi 0.1, so that it lies within the box) offset from the axis
    so that every mode carries a response."""
    from qpad_amd import configs
    return configs.small(nr=32, nxi=24, max_mode=2, ppc=(2, 4), total_num=1 << 10, sigma_xi=0.1,
                         offset_x=0.05)


def oracle_record():
    import pyoracle
    import twin
    pyoracle.lib()
    d = deck()
    o = twin.oracle_from_deck(d)
    try:
        return twin.record_oracle_step(o, o.grid.nr, o.grid.nxi)
    finally:
        o.close()


if __name__ == "__main__":
    rec = oracle_record()
    np.savez_compressed(FIXTURE, **rec)
    print(f"wrote {FIXTURE}: iterations {rec['iters'].tolist()}, "
          f"{len(rec['sp_id'])} plasma and {len(rec['beam_x1'])} beam particles")
