"""Write the golden records of small decks (one or two 3D steps), from the oracle.

QPAD ships no golden vectors for this path (its test programs hold no
assertions), so each fixture is the oracle's own output on a fixed deck,
committed after the oracle was pinned against mathematics.  It serves two
checks: the oracle must keep reproducing it (a change to the oracle that moves
its answer fails a CPU test), and the HIP library must match it on a GPU.

    python tests/golden/make_golden.py      # rewrites every record (*.npz)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]


def deck_small():
    """Nr 32, 24 slices, modes 0..2, electrons (8 particles per cell), a
    1024-particle beam (sigma_xi 0.1, so that it lies within the box) offset
    from the axis so that every mode carries a response."""
    from qpad_amd import configs
    return configs.small(nr=32, nxi=24, max_mode=2, ppc=(2, 4), total_num=1 << 10, sigma_xi=0.1,
                         offset_x=0.05)


def deck_ions_m4():
    """Nr 24, 20 slices, modes 0..4, electrons and ions (neither neutralised
    by a background), the same kind of beam."""
    from qpad_amd import configs
    return configs.small(nr=24, nxi=20, max_mode=4, ppc=(2, 4), total_num=1 << 10, sigma_xi=0.1,
                         offset_x=0.05, ions=True)


def deck_unequal_species():
    """Nr 24, 20 slices, modes 0..2, electrons at 3 x 5 and ions at 1 x 2
    particles per cell (neither neutralised): the species differ in count,
    so the all-species launches run rows past the smaller species' end."""
    from qpad_amd import configs
    return configs.small(nr=24, nxi=20, max_mode=2, ppc=(3, 5), total_num=1 << 10, sigma_xi=0.1,
                         offset_x=0.05, ions=True, ion_ppc=(1, 2))


DECKS = {"step_small": deck_small, "step_ions_m4": deck_ions_m4, "two_steps_small": deck_small,
         "step_unequal_species": deck_unequal_species}
# 3D steps per record (default 1): the second step of two_steps_small starts
# from the beam the first one pushed and from the re-injected plasma
STEPS = {"two_steps_small": 2}


def fixture_path(name):
    return os.path.join(HERE, name + ".npz")


def oracle_record(name, deck=None):
    """The oracle's record of deck `name` (or of `deck`)."""
    import pyoracle
    import twin
    pyoracle.lib()
    o = twin.oracle_from_deck(DECKS[name]() if deck is None else deck)
    try:
        return twin.record_oracle_step(o, o.grid.nr, o.grid.nxi, STEPS.get(name, 1))
    finally:
        o.close()


def load(name):
    with np.load(fixture_path(name)) as z:
        return {k: z[k] for k in z.files}


if __name__ == "__main__":
    for name in (sys.argv[1:] or DECKS):
        rec = oracle_record(name)
        np.savez_compressed(fixture_path(name), **rec)
        nsp = sum(1 for k in rec if k.endswith("_id"))
        print(f"wrote {fixture_path(name)}: iterations {rec['iters'].tolist()}, {nsp} species, "
              f"{len(rec['beam_x1'])} beam particles")
