"""Batched instances (qpad_batch_*): several independent simulations of
different Nr, modes and beams stepped in one call, eagerly (one stream per
instance, queued by a host thread per instance or by one polling host
thread) or by replaying each instance's captured slice graph.  Every
instance must match the same simulation stepped alone with qpad_step3d bit
for bit, on every step: the kernels and their order per instance are the
same, only the instances run side by side.  The graph is replayed on three
consecutive steps with the eager step head and tail (and host
synchronisations) between replays."""
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


@pytest.mark.parametrize("graph", ["eager", "graph", "eager_single"])
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
            assert_same(state(sim), st_ref, f"instance {i} step {step} ({graph})")
    batch.close()
    for sim in sims:
        sim.close()


def test_launch_shapes_and_tuner_change_no_result():
    """Every deposit workgroup shape gives the same run bit for bit, and the
    tuner (timed deposits on the live state between steps) changes nothing."""
    d = scan_decks()[1]
    ref = Simulation(d)
    its_ref = [ref.step3d() for _ in range(2)]
    st_ref = state(ref)
    ref.close()
    for cells in (2, 8):
        sim = Simulation(d)
        assert sim.launch_shape(cells) == cells
        its = [sim.step3d() for _ in range(2)]
        assert its == its_ref, (cells, its, its_ref)
        assert_same(state(sim), st_ref, f"dep_cells={cells}")
        sim.close()
    sim = Simulation(d)
    its = [sim.step3d()]
    ms, choice = sim.tune_launch(reps=2)
    assert choice in (2, 4, 8) and all(v > 0 for v in ms.values()), (ms, choice)
    its.append(sim.step3d())
    assert its == its_ref
    assert_same(state(sim), st_ref, "tuned between steps")
    with pytest.raises(Exception, match="dep_cells"):
        sim.launch_shape(3)
    sim.close()


def test_tuned_graph_batch_matches_alone():
    ref = []
    for d in scan_decks():
        sim = Simulation(d)
        ref.append([(sim.step3d(), state(sim)) for _ in range(2)])
        sim.close()
    sims = [Simulation(d) for d in scan_decks()]
    batch = Batch(sims, graph=True)
    for step in range(2):
        if step == 1:
            assert all(c in (2, 4, 8) for c in batch.tune(reps=2))
        its = batch.step3d()
        for i, sim in enumerate(sims):
            assert its[i] == ref[i][step][0]
            assert_same(state(sim), ref[i][step][1], f"tuned batch instance {i} step {step}")
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


def test_seeded_eager_batch_matches_instances_alone():
    """Instances carrying the learned seed (trained for one step from
    plain-solver labels, then seeding frozen) step in an eager batch exactly
    as each steps alone with qpad_step3d; the graph mode refuses them.  The
    slices that ended at iter_max are counted per step."""
    from qpad_amd import SeedNet

    def seeded(d):
        sim = Simulation(d)
        net = SeedNet(sim, seed=5, lr=1e-5)
        net.mode(seed=False, train=True)
        sim.step3d()
        net.mode(seed=True, train=False)
        return sim

    ref = []
    for d in scan_decks():
        sim = seeded(d)
        ref.append([(sim.step3d(), sim.step_iterations(), state(sim)) for _ in range(2)])
        sim.close()
    for mode in ("eager", "eager_single"):
        sims = [seeded(d) for d in scan_decks()]
        batch = Batch(sims, graph=mode)
        for step in range(2):
            its = batch.step3d()
            for i, sim in enumerate(sims):
                it_ref, stats_ref, st_ref = ref[i][step]
                assert its[i] == it_ref and sim.step_iterations() == stats_ref, (mode, i, step)
                assert stats_ref[0] == it_ref and 0 <= stats_ref[1] <= NXI
                assert_same(state(sim), st_ref, f"seeded {mode} batch instance {i} step {step}")
        batch.mode("graph")
        with pytest.raises(Exception, match="seed network"):
            batch.step3d()
        batch.close()
        for sim in sims:
            sim.close()
