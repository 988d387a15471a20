"""Oracle twin of a GPU Simulation: same grid, same species, the GPU's beam,
and helpers to copy the GPU's slice inputs into the oracle and compare."""
import ctypes as C

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
        if getattr(sp, "neutral", False):
            o.add_neutral(_as_oracle_struct(po.NeutralCfg, c))
            continue
        k = o.add_species(po.species_cfg(c.ppc1, c.ppc2, c.num_theta, c.qm, c.mass, c.density,
                                         tuple(c.uth), bool(c.neutralized), c.den_min, c.seed, c.push_type))
        if getattr(sp, "profile", None) is not None:
            oracle_profile(o, k, sp.profile, 0.0)
            o.renew_species(k)
    for b in sim.beam:
        arr = b.get()
        c = b.cfg
        cfg = po.beam_cfg(0, 1.0, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (-1.0, -1.0, -1.0),
                          (1.0, 1.0, 1.0), qm=c.qm, mass=c.mass, evol=bool(c.evol))
        ib = o.add_beam(cfg, npmax=max(1, len(arr["x1"])))
        o.set_beam(ib, arr["x1"], arr["x2"], arr["x3"], arr["p1"], arr["p2"], arr["p3"], arr["q"])
    return o


def _as_oracle_struct(cls, src):
    """Copy a library config struct into the oracle's struct of the same
    layout, field by field (the oracle's fields are a subset)."""
    out = cls()
    for name, typ in cls._fields_:
        v = getattr(src, name)
        if hasattr(typ, "_length_"):
            C.memmove(C.addressof(getattr(out, name)), C.addressof(v), C.sizeof(typ))
        else:
            setattr(out, name, v)
    return out


def oracle_from_deck(deck, L=None):
    """The oracle set up from a qpinput deck on its own, as qpad_amd.Simulation
    sets up the library: same grid, same species and beam seeds, sub-cycling
    when the deck asks for it; the oracle injects the plasma and the beam
    itself.  L: the oracle library to run (default: the built one)."""
    from qpad_amd.input import (beam_from_deck, grid_from_deck, neutral_from_deck, profile_from_deck,
                                species_from_deck, subcycling_from_deck)
    g, _, z0 = grid_from_deck(deck)
    o = po.Oracle(_as_oracle_struct(po.Grid, g), L)
    sc = subcycling_from_deck(deck)
    if sc is not None:
        o.set_subcycling(*sc)
    sim = deck["simulation"]
    seed = int(sim.get("random_seed", 10))
    nsp = sim.get("nspecies", len(deck.get("species", [])))
    nbm = sim.get("nbeams", len(deck.get("beam", [])))
    for k, sp in enumerate(deck.get("species", [])[:nsp]):
        o.add_species(_as_oracle_struct(po.SpeciesCfg, species_from_deck(sp, seed=seed + 1000 * k)))
        if profile_from_deck(sp) is not None:
            oracle_profile(o, k, profile_from_deck(sp), 0.0)
            o.renew_species(k)
    nnt = sim.get("nneutrals", len(deck.get("neutrals", [])))
    for nt in deck.get("neutrals", [])[:nnt]:
        o.add_neutral(_as_oracle_struct(po.NeutralCfg, neutral_from_deck(nt, sim.get("n0", 1.0e16))))
    for k, bm in enumerate(deck.get("beam", [])[:nbm]):
        cfg = beam_from_deck(bm, z0, seed=seed + 7 * (k + 1))
        if type(cfg).__name__ == "FileBeam":
            import h5read
            with h5read.File(cfg.filename) as f:
                raw = {key: f.dataset(key) for key in ("x1", "x2", "x3", "p1", "p2", "p3", "q")}
            a = po.file_beam_place(raw, cfg.beam_center, cfg.file_center, cfg.length_conv, cfg.charge_conv,
                                   g.nr, g.dr, g.nxi, g.dxi)
            ib = o.add_beam(po.beam_cfg(0, 1.0, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (-1.0, -1.0, -1.0),
                                        (1.0, 1.0, 1.0), qm=cfg.qm, mass=cfg.mass, evol=cfg.evol),
                            npmax=max(1, len(a["x1"])))
            o.set_beam(ib, a["x1"], a["x2"], a["x3"], a["p1"], a["p2"], a["p3"], a["q"])
        elif type(cfg).__name__ == "BeamStdCfg":
            o.add_beam_std(_as_oracle_struct(po.BeamStdCfg, cfg))
        elif type(cfg).__name__ == "TabulatedBeam":
            from qpad_amd.profiles import beam_lattice_box, beam_lattice_density
            box = beam_lattice_box(g.nr, g.nxi, g.dr, g.dxi, cfg.cfg)
            den = beam_lattice_density(cfg.profile, cfg.cfg, g.nr, g.dr, g.dxi, box)
            o.add_beam_std_table(_as_oracle_struct(po.BeamStdCfg, cfg.cfg), den, *box)
        else:
            o.add_beam(_as_oracle_struct(po.BeamCfg, cfg))
    return o


def oracle_profile(o, isp, profile, s):
    """species isp's profile at propagation distance s on the oracle, as
    Species2D.set_profile gives it to the library"""
    from qpad_amd.profiles import plasma_lattice
    c = o.species_raw(isp).contents.cfg
    g = o.grid
    x1, x2 = plasma_lattice(g.nr, g.dr, c.ppc1, c.ppc2, c.num_theta)
    o.set_profile(isp, profile.den_perp(x1, x2, s), profile.den_lon(s))


def oracle_profiles_at(o, deck, s):
    """every profiled species of the deck at s (before the oracle's renew)"""
    from qpad_amd.input import profile_from_deck
    nsp = deck["simulation"].get("nspecies", len(deck.get("species", [])))
    for k, sp in enumerate(deck.get("species", [])[:nsp]):
        prof = profile_from_deck(sp)
        if prof is not None:
            oracle_profile(o, k, prof, s)


def laser_envelope(g, nm, a0=1.0, w0=1.0, xi_c=None, length=0.6, m1=0.15, curv=0.3):
    """A synthetic laser envelope history for the laser-path tests (the laser
    solver that evolves QPAD's envelope is not on the path): real and
    imaginary parts [nxi + 3, nm, nr + 2] over slices -1 .. nxi + 1 (slice s
    at xi = s dxi).  Gaussian in r (waist w0, peak a0) with a wavefront phase
    curv r^2 in the imaginary part, sin^2 in xi over 2 length centred at
    xi_c (default: the middle of the box); m = 1 content m1 (r / w0) times
    the m = 0 envelope where the grid has modes."""
    r = (np.arange(g.nr + 2) - 1.0) * g.dr
    xi = np.arange(-1, g.nxi + 2) * g.dxi
    if xi_c is None:
        xi_c = 0.5 * g.nxi * g.dxi
    d = xi - xi_c
    lon = np.where(np.abs(d) < length, np.cos(0.5 * np.pi * d / length) ** 2, 0.0)
    perp = a0 * np.exp(-(r / w0) ** 2)
    ar = np.zeros((g.nxi + 3, nm, g.nr + 2))
    ai = np.zeros_like(ar)
    ar[:, 0, :] = lon[:, None] * (perp * np.cos(curv * r * r))[None, :]
    ai[:, 0, :] = lon[:, None] * (perp * np.sin(curv * r * r))[None, :]
    if nm >= 3:
        ar[:, 1, :] = m1 * (r / w0)[None, :] * ar[:, 0, :]
        ai[:, 2, :] = m1 * (r / w0)[None, :] * ar[:, 0, :]
        ar[:, 2, :] = -0.5 * m1 * (r / w0)[None, :] * ai[:, 0, :]
    return ar, ai


LASER_FIELDS = (("ar", po.LASER_AR), ("ai", po.LASER_AI), ("ar_grad", po.LASER_AR_GRAD),
                ("ai_grad", po.LASER_AI_GRAD), ("chi", po.LASER_CHI))


def compare_laser(sim, o, norm, j, chi=True):
    """the slice's envelope, gradients (and chi with its history slice j)"""
    for name, which in LASER_FIELDS:
        if name == "chi" and not chi:
            continue
        norm.add("laser." + name, sim.fields[name].get(), o.laser(which), j)
    if chi:
        norm.add("h_chi", sim.fields["chi"].hist(j), o.laser(po.LASER_H_CHI)[j], j)


# what a golden record of one 3D step holds: iteration counts per slice, the
# histories below (every slice), every species' plasma particles after the
# last slice (by id) and beam 0 after its push (each coordinate sorted: the
# library keeps the beam sorted by cell)
GOLDEN_HIST = ("psi", "e", "b", "b_spe", "cu", "q_spe", "q_beam")
GOLDEN_SPECIES = ("x1", "x2", "p1", "p2", "p3", "gamma", "psi")
GOLDEN_BEAM = ("x1", "x2", "x3", "p1", "p2", "p3")


def _record_particles(out, species, b):
    for i, sp in enumerate(species):
        sp = particles_by_id({k: np.array(v) for k, v in sp.items()})
        out[f"sp{i}_id"] = sp["id"].astype(np.int32)
        for k in GOLDEN_SPECIES:
            out[f"sp{i}_{k}"] = sp[k]
    for k in GOLDEN_BEAM:
        out["beam_" + k] = np.sort(np.array(b[k]))


def record_oracle_step(o, nr, nxi, nsteps=1):
    """nsteps 3D steps of the oracle (every species, beam 0): the iteration
    counts of every step, the rest of the record from the last one."""
    iters = []
    for i in range(1, nsteps + 1):
        o.step_begin()
        iters += [o.slice(j) for j in range(1, nxi + 1)]
        if i < nsteps:
            o.step_end(i)
    out = {"iters": np.array(iters, np.int32)}
    for n in GOLDEN_HIST:
        out["hist_" + n] = np.array(o.hist(ORA_ID[n])[1:nxi + 1, :, :, :nr + 2])
    species = [{k: np.array(v) for k, v in o.species(i).items()} for i in range(o.nspecies)]
    o.step_end(nsteps)
    _record_particles(out, species, o.beam(0))
    return out


def record_gpu_step(sim, nsteps=1):
    """The same record of nsteps 3D steps of a qpad_amd.Simulation."""
    nxi = sim.grid.nxi
    iters = []
    for i in range(1, nsteps + 1):
        sim.step3d_begin()
        iters += [sim.slice(j) for j in range(1, nxi + 1)]
        if i < nsteps:
            sim.step3d_end()
    out = {"iters": np.array(iters, np.int32)}
    for n in GOLDEN_HIST:
        out["hist_" + n] = np.stack([sim.fields[n].hist(j) for j in range(1, nxi + 1)])
    species = [sp.get() for sp in sim.spe]
    sim.step3d_end()
    _record_particles(out, species, sim.beam[0].get())
    return out


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


def _oracle_capture(o, nspecies, j):
    out = {name: o.field(ORA_ID[name]).copy() for name in FIELDS}
    for name in HIST:
        out["h_" + name] = o.hist(ORA_ID[name])[j].copy()
    for k in range(nspecies):
        r = particles_by_id({n: np.array(v) for n, v in o.species(k).items()})
        for name in ("x1", "x2", "p1", "p2", "p3", "gamma", "psi"):
            out[f"sp{k}.{name}"] = r[name]
    return out


def oracle_self_sensitivity(o, j, nspecies, eps=1e-15, seed=1):
    """How far the oracle's own slice j moves when its input J_perp (cu, the
    predictor the slice starts from) is perturbed by a relative eps (seeded
    normal): slice j is run from the oracle's current state once perturbed,
    then once as it is, which leaves the oracle in the unperturbed state
    after slice j.  Returns (iterations, ErrorNorm of the perturbed run
    against the unperturbed one over the quantities compare_fields and
    compare_species check).  Where the predictor-corrector iteration does not
    contract (deep in a strong wake) this is far above round-off: no port
    can agree with the oracle more closely than the oracle agrees with
    itself."""
    spec = [{n: np.array(v) for n, v in o.species(k).items()} for k in range(nspecies)]
    fields = {fid: o.field(fid).copy() for fid in range(len(po.NCOMP))}

    def restore():
        for k in range(nspecies):
            o.set_species(k, {n: v.copy() for n, v in spec[k].items()})
        for fid, v in fields.items():
            o.field(fid)[...] = v

    restore()
    cu = o.field(po.CU)
    cu *= 1.0 + eps * np.random.default_rng(seed).standard_normal(cu.shape)
    o.slice(j)
    pert = _oracle_capture(o, nspecies, j)
    restore()
    its = o.slice(j)
    base = _oracle_capture(o, nspecies, j)
    norm = ErrorNorm()
    for name, v in base.items():
        norm.add(name, pert[name], v, j)
    return its, norm
