"""Species of unequal particle counts, and a species the cell sort cannot hold.

The current deposit, the charge deposit, the push and the sort each launch
once for all species, with a grid sized by the largest species; rows past a
smaller species' own count must return at once (k_sort.hip, k_plasma.hip).
Here the electrons crowd every radial cell past three deposit chunks (3000
per cell) while the ions hold 2 per cell, so the species differ in particle
count, sort chunks, deposit grid and overflow rows.  Both are compared with
the oracle slice by slice on identical inputs, through the fused slice and
through the fine-grained entry points (species2d%qdp / %amjdp one species at
a time).  QPAD runs any number of species (sim_plasma_class.f03:116-157).

A species whose lattice exceeds what the cell sort can hold is refused by
qpad_species_new with QPAD_E_INVALID, and the context is left as it was:
the refused index does not exist, and the simulation still steps."""
import ctypes as C

import numpy as np
import pytest

import twin
from qpad_amd import Simulation, configs
from qpad_amd import _lib as L
from test_gpu_fine_grained import fine_grained_slice

pytestmark = pytest.mark.gpu

TOL = 1e-10


def mixed_deck():
    return configs.small(nr=16, nxi=24, max_mode=2, ppc=(25, 120), total_num=1 << 12, n_peak=6.0,
                         sigma_xi=0.1, offset_x=0.03, ions=True, ion_ppc=(1, 2))


@pytest.mark.parametrize("path", ["fused", "fine_grained"])
def test_unequal_species_match_oracle(path):
    sim = Simulation(mixed_deck())
    ne, ni = sim.spe[0].count(), sim.spe[1].count()
    print(f"electrons {ne}, ions {ni} particles")
    assert ne == 16 * 3000 and ni == 16 * 2
    o = twin.make_twin(sim)
    sim.step3d_begin()
    o.step_begin()
    norm = twin.ErrorNorm()
    # slices ahead of, inside and behind the short beam (centred at slice 10)
    for j in range(1, 15):
        twin.sync_slice_inputs(sim, o, j)
        it_g = sim.slice(j) if path == "fused" else fine_grained_slice(sim, j)
        it_o = o.slice(j)
        assert it_g == it_o, f"slice {j}: {it_g} iterations (GPU) vs {it_o} (oracle)"
        names = twin.FIELDS if path == "fused" else [n for n in twin.FIELDS if n != "b_beam"]
        for name in names:
            norm.add(name, sim.fields[name].get(), o.field(twin.ORA_ID[name]), j)
        if path == "fused":
            for name in twin.HIST:
                norm.add("h_" + name, sim.fields[name].hist(j), o.hist(twin.ORA_ID[name])[j], j)
        twin.compare_species(sim, o, norm, j)
    sim.close()
    print({n: f"{e:.1e}" for n, e in sorted(norm.errors().items())})
    bad = norm.bad(TOL)
    assert not bad, f"beyond {TOL}: {bad}"


def _species_cfg(ppc1, ppc2, nth):
    from qpad_amd.input import ctypes_d3
    return L.SpeciesCfg(ppc1, ppc2, nth, -1.0, 1.0, 1.0, ctypes_d3([0.0, 0.0, 0.0]), 1, 1e-10, 0,
                        L.PUSH2_ROBUST)


def test_species_over_sort_limit_is_refused_and_context_survives():
    # nr 4096: 4096 x 8 x 32 = 1,048,576 lattice particles is the most the
    # cell sort holds (npmax = twice the lattice); 4096 x 9 x 32 is refused
    deck = configs.small(nr=4096, nxi=8, max_mode=1, ppc=(1, 2), total_num=1 << 10, rmax=12.0)
    sim = Simulation(deck)
    lib, ctx = sim.lib, sim.ctx
    sid = C.c_int32(-7)
    rc = lib.qpad_species_new(ctx, C.byref(_species_cfg(9, 1, 32)), C.byref(sid))
    assert rc == L.E_INVALID, rc
    assert b"too many particles" in lib.qpad_last_error()
    assert sid.value == -7
    n = C.c_int64(0)
    assert lib.qpad_species_count(ctx, 1, C.byref(n)) == L.E_INVALID
    # the largest lattice the sort holds is accepted and becomes species 1
    rc = lib.qpad_species_new(ctx, C.byref(_species_cfg(8, 1, 32)), C.byref(sid))
    assert rc == 0, lib.qpad_last_error()
    assert sid.value == 1
    assert lib.qpad_species_count(ctx, 1, C.byref(n)) == 0 and n.value == 4096 * 8 * 32
    # the context still steps with both species
    tot = C.c_int32(0)
    assert lib.qpad_step3d(ctx, C.byref(tot)) == 0, lib.qpad_last_error()
    assert 8 <= tot.value <= 8 * sim.grid.iter_max
    sim.close()
