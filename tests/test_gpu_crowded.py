"""Parity where radial cells hold more particles than one deposit wave takes.

A wave deposits at most 1024 particles of a cell; the further chunks of a
crowded cell are deposited by overflow waves into overflow partial rows that
every consumer adds in row order (the charge deposit and the slice sources,
the current deposit and the iteration sources, and the fine-grained combine
kernels).  The small C-style decks elsewhere never crowd a cell that much, so
this deck puts 3000 particles in every radial cell (three chunks: 1024, 1024,
952) and compares slices with the oracle on identical inputs, through the
fused slice and through the fine-grained entry points."""
import numpy as np
import pytest

import twin
from qpad_amd import Simulation, configs
from test_gpu_fine_grained import fine_grained_slice

pytestmark = pytest.mark.gpu

TOL = 1e-10
CHUNK = 1024


def crowded_deck():
    # 25 radial x 120 azimuthal particles per cell = 3000 per cell; a short
    # beam (sigma_xi 0.1) centred at slice 10 of 24, so that the compared
    # slices behind it carry a real plasma response
    return configs.small(nr=16, nxi=24, max_mode=2, ppc=(25, 120), total_num=1 << 12, n_peak=6.0,
                         sigma_xi=0.1, offset_x=0.03)


def max_cell_count(sim):
    p = sim.spe[0].get()
    g = sim.grid
    key = np.floor(np.hypot(p["x1"], p["x2"]) / g.dr).astype(np.int64)
    key = key[key < g.nr]
    return np.bincount(key, minlength=g.nr).max()


@pytest.mark.parametrize("path", ["fused", "fine_grained"])
def test_crowded_cells_match_oracle(path):
    sim = Simulation(crowded_deck())
    assert max_cell_count(sim) > 2 * CHUNK, "the deck must crowd cells past two chunks"
    o = twin.make_twin(sim)
    sim.step3d_begin()
    o.step_begin()
    j0, n = 11, 3
    for j in range(1, j0):
        sim.slice(j)
    norm = twin.ErrorNorm()
    for j in range(j0, j0 + n):
        twin.sync_slice_inputs(sim, o, j)
        it_g = sim.slice(j) if path == "fused" else fine_grained_slice(sim, j)
        it_o = o.slice(j)
        assert it_g == it_o, f"slice {j}: {it_g} iterations (GPU) vs {it_o} (oracle)"
        if path == "fused":
            twin.compare_fields(sim, o, norm, j)
        else:
            # the procedure-by-procedure driver keeps no histories (QPAD's
            # copy_slice calls are its own)
            for name in twin.FIELDS:
                norm.add(name, sim.fields[name].get(), o.field(twin.ORA_ID[name]), j)
        twin.compare_species(sim, o, norm, j)
    # the response is real (not round-off noise): psi is O(0.1) behind the beam
    assert norm.scale["psi"] > 1e-3, norm.scale["psi"]
    sim.close()
    bad = norm.bad(TOL)
    assert not bad, f"{path}: beyond {TOL}: {bad}"


def test_crowded_rerun_is_bit_identical():
    """The overflow rows are added in a fixed order: two runs agree bit for bit."""
    def run():
        sim = Simulation(crowded_deck())
        sim.step3d()
        out = {n: sim.fields[n].get() for n in ("psi", "e", "b", "b_spe", "cu", "q_spe")}
        out.update({"sp." + k: v for k, v in sim.spe[0].get().items()})
        sim.close()
        return out
    a, b = run(), run()
    for k in a:
        assert np.array_equal(a[k], b[k]), f"{k} differs between runs"
