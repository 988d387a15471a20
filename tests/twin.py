"""Oracle twin of a GPU Simulation: same grid, same species, the GPU's beam,
and helpers to copy the GPU's slice inputs into the oracle and compare."""
import numpy as np

import pyoracle as po
from conftest import rel_err

FIELDS = ["psi", "e", "e_spe", "b", "b_spe", "b_beam", "cu", "acu", "amu", "dcu", "q_spe"]
HIST = ["psi", "e", "b", "b_spe", "e_spe", "cu", "q_spe"]
ORA_ID = {n: i for i, n in enumerate(po.FIELD_NAMES)}
# fields are O(1e-2..1) in normalised units; below this scale a field is
# round-off noise and the comparison becomes absolute (tol * FLOOR)
FLOOR = 1e-6


def make_twin(sim):
    g = sim.grid
    og = po.Grid(g.nr, g.nxi, g.max_mode, g.dr, g.dxi, g.relax_fac, g.bnd, g.iter_max,
                 g.iter_reltol, g.iter_abstol, g.dt)
    o = po.Oracle(og)
    for sp in sim.spe:
        c = sp.cfg
        o.add_species(po.species_cfg(c.ppc1, c.ppc2, c.num_theta, c.qm, c.mass, c.density,
                                     tuple(c.uth), bool(c.neutralized), c.den_min, c.seed))
    for b in sim.beam:
        arr = b.get()
        c = b.cfg
        cfg = po.beam_cfg(0, 1.0, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (-1.0, -1.0, -1.0),
                          (1.0, 1.0, 1.0), qm=c.qm, mass=c.mass, evol=bool(c.evol))
        ib = o.add_beam(cfg, npmax=max(1, len(arr["x1"])))
        o.set_beam(ib, arr["x1"], arr["x2"], arr["x3"], arr["p1"], arr["p2"], arr["p3"], arr["q"])
    return o


def sync_slice_inputs(sim, o, j):
    """Copy what slice j depends on: particles, qn, cu, b_spe, q_beam(j)."""
    for k, sp in enumerate(sim.spe):
        o.set_species(k, sp.get())
        _, qn = o.species_q_own(k)
        qn[...] = sp.qn()
    for name in ("cu", "b_spe"):
        o.field(ORA_ID[name])[...] = sim.fields[name].get()
    o.hist(po.Q_BEAM)[j] = sim.fields["q_beam"].hist(j)


def particles_by_id(arrs):
    order = np.argsort(arrs["id"], kind="stable")
    return {k: v[order] for k, v in arrs.items()}


class ErrorNorm:
    """Norm-wise relative error over a whole run: for each quantity,
    max over slices of max |gpu - oracle| divided by max over slices of
    max |oracle|.  Ahead of the beam the plasma response is round-off noise;
    measured against the quantity's own scale over the run, that noise counts
    as the absolute error it is."""

    def __init__(self):
        self.diff = {}
        self.scale = {}
        self.where = {}

    def add(self, name, a, b, j):
        a = np.asarray(a, dtype=np.float64)
        b = np.asarray(b, dtype=np.float64)
        d = float(np.max(np.abs(a - b))) if b.size else 0.0
        s = float(np.max(np.abs(b))) if b.size else 0.0
        if d > self.diff.get(name, -1.0):
            self.diff[name] = d
            self.where[name] = j
        self.scale[name] = max(self.scale.get(name, 0.0), s)

    def errors(self):
        return {n: (self.diff[n] / self.scale[n] if self.scale[n] > 0 else self.diff[n])
                for n in self.diff}

    def bad(self, tol):
        return {n: (e, self.where[n]) for n, e in self.errors().items() if not e <= tol}


def compare_species(sim, o, norm, j, keys=("x1", "x2", "p1", "p2", "p3", "gamma", "psi")):
    for k, sp in enumerate(sim.spe):
        g = particles_by_id(sp.get())
        r = particles_by_id({n: np.array(v) for n, v in o.species(k).items()})
        assert np.array_equal(g["id"], r["id"]), f"slice {j}: particle sets differ"
        for name in keys:
            norm.add(f"sp{k}.{name}", g[name], r[name], j)


def compare_fields(sim, o, norm, j, names=FIELDS):
    for name in names:
        norm.add(name, sim.fields[name].get(), o.field(ORA_ID[name]), j)
    for name in HIST:
        norm.add("h_" + name, sim.fields[name].hist(j), o.hist(ORA_ID[name])[j], j)
