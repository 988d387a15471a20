"""The boundary from QPAD's side on the GPU: a Fortran program
(qpad_amd/fortran/qpad_slice_driver.f03) drives one 3D step through
qpad_hip_mod one procedure at a time, in the order of QPAD's
simulation%run (simulation_class.f03:294-469), with QPAD's host arithmetic
(zeroing, cbq, the J_perp predictor) on its own arrays.  Its iterations,
slice fields and particles equal the fused qpad_slice path's."""
import os

import numpy as np
import pytest

import fortran_driver
import twin
from qpad_amd import Simulation, configs

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(fortran_driver.DRIVER), reason="Fortran driver not built")]

TOL = 1e-10


@pytest.mark.parametrize("max_mode,ions,push", [(1, False, "robust"), (2, True, "robust"), (2, True, "standard")])
def test_fortran_driver_step_matches_fused_path(tmp_path, max_mode, ions, push):
    deck = configs.small(nr=48, nxi=40, max_mode=max_mode, ppc=(2, 4), total_num=1 << 12, ions=ions,
                         offset_x=0.03, push_type=push)
    print("fortran driver: max_mode", max_mode, "ions", ions, flush=True)
    g, iters, fields, parts = fortran_driver.run(deck, str(tmp_path))
    sim = Simulation(deck)
    for k in fortran_driver.GRID_KEYS:
        assert g[k] == getattr(sim.grid, k), k           # the parameters crossed exactly
    sim.step3d_begin()
    ref = np.array([sim.slice(j) for j in range(1, sim.grid.nxi + 1)])
    assert np.array_equal(iters, ref), (iters, ref)
    norm = twin.ErrorNorm()
    j = sim.grid.nxi
    for name in fortran_driver.DUMP:
        norm.add(name, fields[name], sim.fields[name].get(), j)
    mine = twin.particles_by_id(sim.spe[0].get())
    theirs = twin.particles_by_id(parts)
    for name in ("x1", "x2", "p1", "p2", "p3"):
        norm.add("sp0." + name, theirs[name], mine[name], j)
    sim.close()
    print({n: f"{e:.1e}" for n, e in sorted(norm.errors().items())})
    bad = norm.bad(TOL)
    assert not bad, f"beyond {TOL}: {bad}"
