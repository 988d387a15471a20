"""The standard plasma pusher (species push_type "standard") on the GPU
against the oracle: species2d%interp_psi after the psi solve
(simulation_class.f03:357-359, part2d_class.f03:2264-2305), the standard
current deposit (amjdeposit_std_part2d, :478-744) and push
(push_u_std_part2d, :1790-1877), alone and beside a robust species, slice by
slice on identical inputs, then a free-running 3D step with the beam push."""
import numpy as np
import pytest

import twin
from conftest import rel_err
from qpad_amd import Simulation, configs

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.mark.parametrize("max_mode,ions", [(1, False), (2, True)])
def test_standard_pusher_slices_match_oracle(max_mode, ions):
    deck = configs.small(nr=48, nxi=40, max_mode=max_mode, ppc=(2, 4), total_num=1 << 12, ions=ions,
                         ion_ppc=(1, 2), offset_x=0.03, push_type="standard", ion_push_type="robust")
    print("standard pusher: max_mode", max_mode, "ions", ions, flush=True)
    sim = Simulation(deck)
    assert [sp.cfg.push_type for sp in sim.spe] == ([0, 1] if ions else [0])
    o = twin.make_twin(sim)
    sim.step3d_begin()
    o.step_begin()
    norm = twin.ErrorNorm()
    for j in range(1, sim.grid.nxi + 1):
        twin.sync_slice_inputs(sim, o, j)
        it_g, it_o = sim.slice(j), o.slice(j)
        assert it_g == it_o, (j, it_g, it_o)
        twin.compare_fields(sim, o, norm, j)
        twin.compare_species(sim, o, norm, j)
    print({n: f"{e:.1e}" for n, e in sorted(norm.errors().items())})
    bad = norm.bad(TOL)
    assert not bad, f"beyond {TOL}: {bad}"
    sim.close()
    o.close()


def test_standard_pusher_step_and_beam_push_match_oracle():
    """a whole 3D step from the deck alone (the oracle injects its own plasma
    and beam), histories compared norm-wise, then the beam after its push"""
    deck = configs.small(nr=40, nxi=32, max_mode=1, ppc=(2, 4), total_num=1 << 12, offset_x=0.03,
                         push_type="standard")
    sim = Simulation(deck)
    o = twin.oracle_from_deck(deck)
    sim.step3d()
    o.step_begin()
    for j in range(1, sim.grid.nxi + 1):
        o.slice(j)
    o.step_end(1)
    for name in ("psi", "e", "b", "b_spe", "cu", "q_spe"):
        g = np.stack([sim.fields[name].hist(j) for j in range(1, sim.grid.nxi + 1)])
        r = o.hist(twin.ORA_ID[name])[1:sim.grid.nxi + 1]
        e = rel_err(g, r)
        assert e <= TOL, (name, e)
    gb, ob = sim.beam[0].get(), o.beam(0)
    assert len(gb["x1"]) == len(ob["x1"])
    for k in ("x1", "x2", "x3", "p1", "p2", "p3"):
        assert rel_err(np.sort(gb[k]), np.sort(ob[k])) <= 1e-9, k
    sim.close()
    o.close()
