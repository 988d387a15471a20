"""Batched instances (qpad_batch_*): several independent simulations of
different Nr, modes and beams stepped in one call, eagerly (slices queued
round-robin, one stream per instance) or by replaying each instance's
captured slice graph.  Every instance must match the same simulation stepped
alone with qpad_step3d bit for bit, on every step: the kernels and their
order per instance are the same, only the instances run side by side.  The
graph is replayed on three consecutive steps with the eager step head and
tail (and host synchronisations) between replays."""
import numpy as np
import pytest

from qpad_amd import Batch, Simulation, configs

pytestmark = pytest.mark.gpu

NXI = 32


def scan_decks():
    return [configs.small(nr=48, nxi=NXI, max_mode=1, ppc=(2, 4), total_num=1 << 12),
            configs.small(nr=96, nxi=NXI, max_mode=2, ppc=(2, 4), total_num=1 << 12, offset_x=0.03,
                          n_peak=6.0),
            configs.small(nr=64, nxi=NXI, max_mode=4, ppc=(2, 4), total_num=1 << 12, offset_x=0.05,
                          ions=True)]


def state(sim):
    out = {n: np.stack([sim.fields[n].hist(j) for j in range(1, sim.grid.nxi + 1)])
           for n in ("psi", "b_spe", "e", "cu")}
    for k, sp in enumerate(sim.spe):
        out.update({f"sp{k}.{n}": v for n, v in sp.get().items()})
    out.update({"beam." + n: v for n, v in sim.beam[0].get().items()})
    return out


def assert_same(a, b, what):
    for k in a:
        assert np.array_equal(a[k], b[k]), f"{what}: {k} differs"


@pytest.mark.parametrize("graph", [False, True])
def test_batch_matches_instances_stepped_alone(graph):
    nsteps = 3
    ref = []
    for d in scan_decks():
        sim = Simulation(d)
        steps = []
        for _ in range(nsteps):
            it = sim.step3d()
            steps.append((it, state(sim)))
        ref.append(steps)
        sim.close()
    sims = [Simulation(d) for d in scan_decks()]
    batch = Batch(sims, graph=graph)
    for step in range(nsteps):
        its = batch.step3d()
        for i, sim in enumerate(sims):
            it_ref, st_ref = ref[i][step]
            assert its[i] == it_ref, f"instance {i} step {step}: {its[i]} vs {it_ref} iterations"
            assert_same(state(sim), st_ref, f"instance {i} step {step} ({'graph' if graph else 'eager'})")
    batch.close()
    for sim in sims:
        sim.close()


def test_graph_mode_needs_even_nxi():
    deck = configs.small(nr=48, nxi=33, max_mode=1, ppc=(2, 4), total_num=1 << 10)
    sim = Simulation(deck)
    batch = Batch([sim], graph=True)
    with pytest.raises(Exception, match="even nxi"):
        batch.step3d()
    batch.close()
    sim.close()
