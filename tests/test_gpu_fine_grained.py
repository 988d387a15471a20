"""Parity of the fine-grained entry points, which the Fortran side would call
one procedure at a time.  A slice is driven here in the order of QPAD's slice
body (simulation_class.f03:342-469) through species2d%qdp / %amjdp / %push,
field_psi%solve, field_e%solve, field_b%solve, field_djdxi%solve, add_f1
and convergence_tester.  QPAD's own array arithmetic (zeroing, cbq, the
J_perp predictor) is done on host copies, as the Fortran side would.  Each
slice is compared with the oracle on identical inputs."""
import numpy as np
import pytest

import twin
from qpad_amd import Simulation, configs

pytestmark = pytest.mark.gpu

TOL = 1e-10


def fine_grained_slice(sim, j):
    f = sim.fields
    g = sim.grid
    lib, ctx = sim.lib, sim.ctx
    f["b_beam"].solve(f["q_beam"], slice_index=j)          # q_beam%copy_slice + field_b%solve(rho)
    f["q_spe"].set(np.zeros(f["q_spe"].shape()))
    for sp in sim.spe:
        sp.qdp(f["q_spe"])
    f["psi"].solve(f["q_spe"])
    for sp in sim.spe:
        sp.interp_psi(f["psi"])                              # standard pusher only
    f["b_spe"].solve(f["cu"])                                # Bz from the J predictor
    iters = 0
    for l in range(1, g.iter_max + 1):
        iters = l
        sim.convergence_tester("record")
        assert lib.qpad_field_add_b(ctx) == 0                # b = b_spe + b_beam
        f["e"].solve(f["cu"])
        f["e"].solve(f["b"], f["psi"])
        for name in ("cu", "acu", "amu"):
            f[name].set(np.zeros(f[name].shape()))
        for sp in sim.spe:
            sp.amjdp(f["e"], f["b"], f["cu"], f["amu"], f["acu"], g.dxi)
        f["dcu"].solve(f["acu"], f["amu"])
        f["b_spe"].solve(f["dcu"], f["cu"])
        f["b_spe"].solve(f["cu"])
        rel, ab = sim.convergence_tester("compare")
        if rel < g.iter_reltol or ab < g.iter_abstol:
            break
    # species2d%cbq and the epilogue (simulation_class.f03:401-456)
    cu = f["cu"].get()
    q = f["q_spe"].get()
    q[0] = q[0] + cu[2]
    f["q_spe"].set(q)
    assert lib.qpad_field_add_b(ctx) == 0
    f["e_spe"].solve(f["b_spe"], f["psi"])
    f["e"].solve(f["cu"])
    f["e"].solve(f["b"], f["psi"])
    dcu = f["dcu"].get()
    dcu = g.dxi * dcu
    f["dcu"].set(dcu)
    cu = f["cu"].get()
    cu[0:2] = cu[0:2] + dcu
    f["cu"].set(cu)
    for sp in sim.spe:
        sp.push(g.dxi)
    return iters


@pytest.mark.parametrize("max_mode,ions,push", [(1, False, "robust"), (2, True, "robust"), (2, True, "standard")])
def test_fine_grained_slices_match_oracle(max_mode, ions, push):
    deck = configs.small(nr=40, nxi=48, max_mode=max_mode, ppc=(2, 4), total_num=1 << 12, ions=ions,
                         offset_x=0.03, push_type=push)
    sim = Simulation(deck)
    o = twin.make_twin(sim)
    sim.step3d_begin()
    o.step_begin()
    norm = twin.ErrorNorm()
    fields = ["psi", "e", "e_spe", "b", "b_spe", "b_beam", "cu", "acu", "amu", "dcu", "q_spe"]
    for j in range(1, sim.grid.nxi + 1):
        twin.sync_slice_inputs(sim, o, j)
        it_g = fine_grained_slice(sim, j)
        it_o = o.slice(j)
        assert it_g == it_o, f"slice {j}: {it_g} iterations (GPU) vs {it_o} (oracle)"
        for name in fields:
            norm.add(name, sim.fields[name].get(), o.field(twin.ORA_ID[name]), j)
        twin.compare_species(sim, o, norm, j)
    sim.close()
    bad = norm.bad(TOL)
    assert not bad, f"beyond {TOL}: {bad}"


def test_fine_grained_and_fused_slices_agree():
    """The fused qpad_slice and the procedure-by-procedure path give the same
    slice (both are checked against the oracle above; this pins them to each
    other on a longer run)."""
    deck = configs.small(nr=48, nxi=48, max_mode=2, ppc=(2, 4), total_num=1 << 12, offset_x=0.03)
    a = Simulation(deck)
    b = Simulation(deck)
    a.step3d_begin()
    b.step3d_begin()
    norm = twin.ErrorNorm()
    for j in range(1, a.grid.nxi + 1):
        it_a = a.slice(j)
        it_b = fine_grained_slice(b, j)
        assert it_a == it_b
        for name in ("psi", "e", "b", "b_spe", "cu", "q_spe"):
            norm.add(name, b.fields[name].get(), a.fields[name].get(), j)
        pa = twin.particles_by_id(a.spe[0].get())
        pb = twin.particles_by_id(b.spe[0].get())
        for name in ("x1", "x2", "p1", "p2", "p3"):
            norm.add("sp0." + name, pb[name], pa[name], j)
    a.close()
    b.close()
    bad = norm.bad(TOL)
    assert not bad, f"beyond {TOL}: {bad}"
