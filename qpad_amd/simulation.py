"""Host-side mirror of QPAD's simulation objects over libqpad_hip.

The classes keep the reference's names and call sequence
(source/simulation_class.f03, species/species2d_class.f03,
fields/field_{psi,e,b,src}_class.f03, beam/beam3d_class.f03); every method
is one C ABI call, and all state stays in device memory.  Errors raise
QpadError (the reference's write_err stops the run).
"""
import ctypes as C

import numpy as np

from . import _lib as L
from .input import (FileBeam, TabulatedBeam, beam_from_deck, grid_from_deck, neutral_from_deck, place_file_beam,
                    profile_from_deck, species_from_deck, subcycling_from_deck)


class Field:
    """Handle on one slice field of the context (field_class's rf_re/rf_im)."""

    def __init__(self, sim, name):
        self.sim = sim
        self.name = name
        self.fid = L.FIELD_IDS[name]
        self.ncomp = L.NCOMP[self.fid]

    def shape(self):
        return (self.ncomp, self.sim.nm, self.sim.grid.nr + 2)

    def get(self):
        """f1 as [ncomp, 2*max_mode+1, nr+2] (k=0 Re m0, 2m-1 Re m, 2m Im m)."""
        out = np.zeros(self.shape())
        L.check(self.sim.lib.qpad_get_field(self.sim.ctx, self.fid, L.dptr(out)))
        return out

    def set(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(self.shape())
        L.check(self.sim.lib.qpad_set_field(self.sim.ctx, self.fid, L.dptr(a)))

    def hist(self, s):
        """f2 slice s (1..nxi, nxi+1 = upper guard slice)."""
        out = np.zeros(self.shape())
        L.check(self.sim.lib.qpad_get_hist(self.sim.ctx, self.fid, s, L.dptr(out)))
        return out

    def set_hist(self, s, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(self.shape())
        L.check(self.sim.lib.qpad_set_hist(self.sim.ctx, self.fid, s, L.dptr(a)))


class FieldPsi(Field):
    def solve(self, q):
        """field_psi%solve(q_spe) (field_psi_class.f03:218-254)."""
        if q.name != "q_spe":
            raise L.QpadError("field_psi%solve takes q_spe")
        L.check(self.sim.lib.qpad_field_solve_psi(self.sim.ctx))


class FieldE(Field):
    def solve(self, *args):
        """solve(jay): Ez (field_e_class.f03:298-334); solve(b, psi): E_perp
        (field_e_class.f03:412-514)."""
        if len(args) == 1:
            L.check(self.sim.lib.qpad_field_solve_ez(self.sim.ctx, self.fid, args[0].fid))
        elif len(args) == 2:
            if args[1].name != "psi":
                raise L.QpadError("field_e%solve(b, psi) takes psi")
            L.check(self.sim.lib.qpad_field_solve_et(self.sim.ctx, self.fid, args[0].fid))
        else:
            raise L.QpadError("field_e%solve takes (jay) or (b, psi)")


class FieldB(Field):
    def solve(self, *args, slice_index=None):
        """solve(jay): Bz; solve(q_beam): beam B_perp of slice `slice_index`;
        solve(djdxi, jay): plasma B_perp (field_b_class.f03:760-879)."""
        if len(args) == 1 and args[0].name == "q_beam":
            if self.name != "b_beam" or slice_index is None:
                raise L.QpadError("beam B_perp is solved into b_beam for a given slice")
            L.check(self.sim.lib.qpad_field_solve_bt_beam(self.sim.ctx, int(slice_index)))
        elif len(args) == 1:
            L.check(self.sim.lib.qpad_field_solve_bz(self.sim.ctx, self.fid, args[0].fid))
        elif len(args) == 2:
            if self.name != "b_spe" or args[0].name != "dcu" or args[1].name != "cu":
                raise L.QpadError("the B_perp iteration solves b_spe from (dcu, cu)")
            L.check(self.sim.lib.qpad_field_solve_bt_iter(self.sim.ctx))
        else:
            raise L.QpadError("bad field_b%solve arguments")


class FieldDjdxi(Field):
    def solve(self, acu, amu):
        """field_djdxi%solve(acu, amu) (field_src_class.f03:273-405)."""
        if self.name != "dcu" or acu.name != "acu" or amu.name != "amu":
            raise L.QpadError("dcu%solve takes (acu, amu)")
        L.check(self.sim.lib.qpad_field_solve_djdxi(self.sim.ctx))


class Species2D:
    """species2d (species2d_class.f03:24-54): standard, robust, standard_pgc
    or robust_pgc pusher."""

    def __init__(self, sim, cfg, profile=None):
        self.sim = sim
        sid = C.c_int32(0)
        L.check(sim.lib.qpad_species_new(sim.ctx, C.byref(cfg), C.byref(sid)))
        self.id = sid.value
        self.cfg = cfg
        self.profile = None
        if profile is not None:
            self.set_profile(profile, 0.0)
            self.renew()

    def set_profile(self, profile, s):
        """the density of the species' next injections: a PlasmaProfile
        (qpad_amd.profiles) evaluated at the injection lattice and at
        propagation distance s (inject_fdist2d, fdist2d_class.f03:312-336);
        None: uniform"""
        self.profile = profile
        if profile is None:
            L.check(self.sim.lib.qpad_species_set_profile(self.sim.ctx, self.id, None, 1.0))
            return
        from .profiles import plasma_lattice
        g, c = self.sim.grid, self.cfg
        x1, x2 = plasma_lattice(g.nr, g.dr, c.ppc1, c.ppc2, c.num_theta)
        den = np.ascontiguousarray(profile.den_perp(x1, x2, s), dtype=np.float64)
        L.check(self.sim.lib.qpad_species_set_profile(self.sim.ctx, self.id, L.dptr(den),
                                                      float(profile.den_lon(s))))

    def renew(self):
        L.check(self.sim.lib.qpad_species_renew(self.sim.ctx, self.id))

    def qdp(self, q):
        if q.name != "q_spe":
            raise L.QpadError("species2d%qdp deposits into q_spe")
        L.check(self.sim.lib.qpad_species_qdp(self.sim.ctx, self.id))

    def amjdp(self, ef, bf, cu, amu, acu, dxi):
        if (ef.name, bf.name, cu.name, amu.name, acu.name) != ("e", "b", "cu", "amu", "acu"):
            raise L.QpadError("species2d%amjdp(e, b, cu, amu, acu, dxi)")
        L.check(self.sim.lib.qpad_species_amjdp(self.sim.ctx, self.id, float(dxi)))

    def interp_psi(self, psi):
        """species2d%interp_psi (species2d_class.f03:439-452): psi at the
        particles of a standard-pusher species (no effect on a robust one)"""
        if psi.name != "psi":
            raise L.QpadError("species2d%interp_psi takes psi")
        L.check(self.sim.lib.qpad_species_interp_psi(self.sim.ctx, self.id))

    def push(self, dxi, exp_fac_clamped=None):
        """push_u + push_x (+ update_bound); with exp_fac_clamped, the
        sub-cycled push_u, clamp_exp_fac, push_x (part2d_subcyc_class.f03:
        46-66)"""
        if exp_fac_clamped is None:
            L.check(self.sim.lib.qpad_species_push(self.sim.ctx, self.id, float(dxi)))
        else:
            L.check(self.sim.lib.qpad_species_push_clamped(self.sim.ctx, self.id, float(dxi),
                                                           float(exp_fac_clamped)))

    def count(self):
        n = C.c_int64(0)
        L.check(self.sim.lib.qpad_species_count(self.sim.ctx, self.id, C.byref(n)))
        return n.value

    def get(self):
        n = self.count()
        out = {k: np.zeros(n) for k in ("x1", "x2", "p1", "p2", "p3", "q", "gamma", "psi")}
        out["id"] = np.zeros(n, dtype=np.int32)
        L.check(self.sim.lib.qpad_get_species(
            self.sim.ctx, self.id, *[L.dptr(out[k]) for k in ("x1", "x2", "p1", "p2", "p3", "q",
                                                               "gamma", "psi")], L.iptr(out["id"])))
        return out

    def set(self, arrs):
        n = len(arrs["x1"])
        a = {k: (np.ascontiguousarray(arrs[k], dtype=np.float64) if k in arrs else None)
             for k in ("x1", "x2", "p1", "p2", "p3", "q", "gamma", "psi")}
        ids = np.ascontiguousarray(arrs["id"], dtype=np.int32) if "id" in arrs else None
        L.check(self.sim.lib.qpad_set_species(
            self.sim.ctx, self.id, n, *[L.dptr(a[k]) for k in ("x1", "x2", "p1", "p2", "p3", "q",
                                                              "gamma", "psi")], L.iptr(ids)))

    def charge_hist(self, s):
        """species2d%q history slice s (qpad_species_charge_history on)"""
        out = np.zeros((1, self.sim.nm, self.sim.grid.nr + 2))
        L.check(self.sim.lib.qpad_get_species_charge(self.sim.ctx, self.id, int(s), L.dptr(out)))
        return out

    def qn(self):
        out = np.zeros((1, self.sim.nm, self.sim.grid.nr + 2))
        L.check(self.sim.lib.qpad_get_species_qn(self.sim.ctx, self.id, L.dptr(out)))
        return out

    def set_qn(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        L.check(self.sim.lib.qpad_set_species_qn(self.sim.ctx, self.id, L.dptr(a)))


class Neutral2D(Species2D):
    """neutral (neutral_class.f03): a gas ionised by the slice's field; its
    particles are the electrons ionisation released so far."""

    def __init__(self, sim, cfg):
        self.sim = sim
        sid = C.c_int32(0)
        L.check(sim.lib.qpad_neutral_new(sim.ctx, C.byref(cfg), C.byref(sid)))
        self.id = sid.value
        self.cfg = cfg
        self.neutral = True

    def ion_deposit(self, q):
        """neutral%ion_deposit(q_spe) (neutral_class.f03:904-930)"""
        if q.name != "q_spe":
            raise L.QpadError("neutral%ion_deposit adds into q_spe")
        L.check(self.sim.lib.qpad_neutral_ion_deposit(self.sim.ctx, self.id))

    def update(self):
        """neutral%update(e, psi, s) (neutral_class.f03:576-598)"""
        L.check(self.sim.lib.qpad_neutral_update(self.sim.ctx, self.id))

    def levels(self):
        """[num_theta, nr, ion_max + 2]: level fractions, neutral residue,
        discrete total ion charge"""
        out = np.zeros((self.cfg.num_theta, self.sim.grid.nr, self.cfg.ion_max + 2))
        L.check(self.sim.lib.qpad_get_neutral_levels(self.sim.ctx, self.id, L.dptr(out)))
        return out

    def ion_density(self, s):
        """rho_ion of history slice s"""
        out = np.zeros((1, self.sim.nm, self.sim.grid.nr + 2))
        L.check(self.sim.lib.qpad_get_ion_density(self.sim.ctx, self.id, int(s), L.dptr(out)))
        return out


class Beam3D:
    """beam3d (beam3d_class.f03:51-52): a lattice ("standard"), random
    Gaussian or file beam, reduced pusher."""

    def __init__(self, sim, cfg):
        self.sim = sim
        bid = C.c_int32(0)
        if isinstance(cfg, FileBeam):
            from . import diag
            g = sim.grid
            a = place_file_beam(diag.read_beam_file(cfg.filename), cfg, g.nr, g.dr, g.nxi, g.dxi)
            L.check(sim.lib.qpad_beam_new_arrays(sim.ctx, cfg.qm, cfg.mass, int(cfg.evol), len(a["x1"]),
                                                 *[L.dptr(a[k]) for k in ("x1", "x2", "x3", "p1", "p2", "p3",
                                                                          "q")], C.byref(bid)))
        elif isinstance(cfg, TabulatedBeam):
            # a lattice beam whose profile the host evaluates over the box of its ranges
            from .profiles import beam_lattice_density
            k0, nk, ni = C.c_int32(0), C.c_int32(0), C.c_int32(0)
            L.check(sim.lib.qpad_beam_lattice_box(sim.ctx, C.byref(cfg.cfg), C.byref(k0), C.byref(nk), C.byref(ni)))
            box = (k0.value, nk.value, ni.value)
            g = sim.grid
            den = beam_lattice_density(cfg.profile, cfg.cfg, g.nr, g.dr, g.dxi, box)
            L.check(sim.lib.qpad_beam_new_std_density(sim.ctx, C.byref(cfg.cfg), L.dptr(den), *box, C.byref(bid)))
            cfg = cfg.cfg
        else:
            new = sim.lib.qpad_beam_new_std if isinstance(cfg, L.BeamStdCfg) else sim.lib.qpad_beam_new
            L.check(new(sim.ctx, C.byref(cfg), C.byref(bid)))
        self.id = bid.value
        self.cfg = cfg

    def count(self):
        n = C.c_int64(0)
        L.check(self.sim.lib.qpad_beam_count(self.sim.ctx, self.id, C.byref(n)))
        return n.value

    def get(self):
        n = self.count()
        out = {k: np.zeros(n) for k in ("x1", "x2", "x3", "p1", "p2", "p3", "q")}
        L.check(self.sim.lib.qpad_get_beam(self.sim.ctx, self.id,
                                           *[L.dptr(out[k]) for k in ("x1", "x2", "x3", "p1", "p2",
                                                                       "p3", "q")]))
        return out

    def set(self, arrs):
        a = [np.ascontiguousarray(arrs[k], dtype=np.float64) for k in ("x1", "x2", "x3", "p1",
                                                                       "p2", "p3", "q")]
        L.check(self.sim.lib.qpad_set_beam(self.sim.ctx, self.id, len(a[0]),
                                           *[L.dptr(x) for x in a]))


class SeedNet:
    """The learned B_perp seed of the predictor-corrector iteration (no
    counterpart in QPAD: it replaces the b_spe(xi-1) seed that QPAD's
    iteration starts from, simulation_class.f03:326,370-375).  A 1D conv
    network over r; BF16 weights and activations, FP32 accumulation, Adam on
    FP32 master weights, all on the GPU (include/qpad_hip.h)."""

    def __init__(self, sim, width=64, layers=4, lr=1e-4, beta1=0.9, beta2=0.999, eps=1e-8, out_scale=0.02,
                 seed=1):
        self.sim = sim
        self.cfg = L.SeedNetCfg(width, layers, lr, beta1, beta2, eps, out_scale, seed)
        L.check(sim.lib.qpad_seednet_create(sim.ctx, C.byref(self.cfg)))

    def info(self):
        i = L.SeedNetInfo()
        L.check(self.sim.lib.qpad_seednet_get_info(self.sim.ctx, C.byref(i)))
        return i

    @property
    def nparams(self):
        return self.info().nparams

    def mode(self, seed, train):
        L.check(self.sim.lib.qpad_seednet_mode(self.sim.ctx, int(bool(seed)), int(bool(train))))

    def get_params(self):
        th = np.zeros(self.nparams, np.float32)
        L.check(self.sim.lib.qpad_seednet_get_params(self.sim.ctx, L.fptr(th), th.size))
        return th

    def set_params(self, th):
        th = np.ascontiguousarray(th, np.float32)
        L.check(self.sim.lib.qpad_seednet_set_params(self.sim.ctx, L.fptr(th), th.size))

    def losses(self, first=0, n=None):
        steps = self.info().steps
        n = steps - first if n is None else n
        out = np.zeros(max(n, 0), np.float32)
        if n > 0:
            L.check(self.sim.lib.qpad_seednet_losses(self.sim.ctx, int(first), int(n), L.fptr(out)))
        return out

    def seed(self, j):
        L.check(self.sim.lib.qpad_seednet_seed(self.sim.ctx, int(j)))

    def learn(self, j):
        L.check(self.sim.lib.qpad_seednet_learn(self.sim.ctx, int(j)))

    def eval(self, x, t=None, grad=False, update=False):
        """y (and with t: loss, gradient) of the network on host inputs x [npos][cin]."""
        inf = self.info()
        x = np.ascontiguousarray(x, np.float32)
        npos = x.shape[0]
        assert x.shape == (npos, inf.cin)
        y = np.zeros((npos, inf.cout), np.float32)
        loss = C.c_float(0.0)
        g = np.zeros(inf.nparams, np.float32) if grad else None
        tt = None if t is None else np.ascontiguousarray(t, np.float32)
        L.check(self.sim.lib.qpad_seednet_eval(self.sim.ctx, L.fptr(x), L.fptr(tt), npos, L.fptr(y),
                                               C.byref(loss) if tt is not None else None, L.fptr(g),
                                               int(bool(update))))
        return y, (loss.value if tt is not None else None), g


class Simulation:
    """simulation (simulation_class.f03:34-69) driven from a qpinput deck."""

    def __init__(self, deck, lib_path=None):
        self.lib = L.load(lib_path)
        self.deck = deck
        self.grid, self.nstep3d, z0 = grid_from_deck(deck)
        self.nm = 2 * self.grid.max_mode + 1
        ctx = C.c_void_p()
        L.check(self.lib.qpad_create(C.byref(self.grid), C.byref(ctx)))
        self.ctx = ctx
        seed = int(deck["simulation"].get("random_seed", 10))
        self.fields = {name: cls(self, name) for name, cls in (
            ("psi", FieldPsi), ("e", FieldE), ("e_spe", FieldE), ("b", FieldB),
            ("b_spe", FieldB), ("b_beam", FieldB), ("cu", Field), ("acu", Field), ("amu", Field),
            ("dcu", FieldDjdxi), ("q_spe", Field), ("q_beam", Field),
            ("ar", Field), ("ai", Field), ("ar_grad", Field), ("ai_grad", Field), ("chi", Field))}
        nsp = deck["simulation"].get("nspecies", len(deck.get("species", [])))
        nbm = deck["simulation"].get("nbeams", len(deck.get("beam", [])))
        self.spe = [Species2D(self, species_from_deck(sp, seed=seed + 1000 * k), profile_from_deck(sp))
                    for k, sp in enumerate(deck.get("species", [])[:nsp])]
        # neutrals after the species, as simulation%new creates them
        nnt = deck["simulation"].get("nneutrals", len(deck.get("neutrals", [])))
        n0 = float(deck["simulation"].get("n0", 1.0e16))
        self.neut = [Neutral2D(self, neutral_from_deck(nt, n0)) for nt in deck.get("neutrals", [])[:nnt]]
        self.spe += self.neut
        self.beam = [Beam3D(self, beam_from_deck(bm, z0, seed=seed + 7 * (k + 1)))
                     for k, bm in enumerate(deck.get("beam", [])[:nbm])]
        sc = subcycling_from_deck(deck)
        if sc is not None:
            self.set_subcycling(*sc)
        self.tstep = 0

    # -- sub-cycling (simulation.algorithm "subcycling", proj_subcyc)
    def set_subcycling(self, exp_fac_max, exp_fac_clamped, dt_2d_min, on=True):
        """simulation_subcyc%new (simulation_subcyc_class.f03:76-88)"""
        L.check(self.lib.qpad_set_subcycling(self.ctx, int(on), float(exp_fac_max), float(exp_fac_clamped),
                                             float(dt_2d_min)))

    def exp_fac_max(self):
        """sim_plasma_subcyc%get_exp_fac_max: max(1, gamma / (gamma - p_z))"""
        out = C.c_double(0.0)
        L.check(self.lib.qpad_exp_fac_max(self.ctx, C.byref(out)))
        return out.value

    def last_subcycles(self):
        """(sub-cycles, their 2D step, the expansion factor) of the last slice"""
        n, d, f = C.c_int32(0), C.c_double(0.0), C.c_double(0.0)
        L.check(self.lib.qpad_last_subcycles(self.ctx, C.byref(n), C.byref(d), C.byref(f)))
        return n.value, d.value, f.value

    # -- laser (PGC pushers)
    def set_laser(self, ar_hist, ai_hist):
        """The laser envelope of every slice -1 .. nxi + 1 (field_laser's
        cfr, cfi f2), real and imaginary part, [nxi + 3, nm, nr + 2] each.
        From then on every slice pushes the PGC species in it and deposits
        chi (fields["chi"], with a history)."""
        shape = (self.grid.nxi + 3, self.nm, self.grid.nr + 2)
        a = np.ascontiguousarray(ar_hist, dtype=np.float64)
        b = np.ascontiguousarray(ai_hist, dtype=np.float64)
        if a.shape != shape or b.shape != shape:
            raise L.QpadError(f"laser envelope history must be {shape}")
        L.check(self.lib.qpad_set_laser(self.ctx, L.dptr(a), L.dptr(b)))

    def laser_slice(self, j):
        """laser_all of slice j (simulation_class.f03:362-367): fields ar,
        ai, ar_grad, ai_grad"""
        L.check(self.lib.qpad_laser_slice(self.ctx, int(j)))

    def deposit_chi(self, j):
        """sim_lasers%deposit_chi(spe, j) (sim_lasers_class.f03:176-200)"""
        L.check(self.lib.qpad_deposit_chi(self.ctx, int(j)))

    # -- fused run loop (simulation_class.f03:226-520)
    def step3d_begin(self):
        L.check(self.lib.qpad_step3d_begin(self.ctx))

    def slice(self, j):
        it = C.c_int32(0)
        L.check(self.lib.qpad_slice(self.ctx, int(j), C.byref(it)))
        return it.value

    def _profiles_at(self, s):
        """the profiled species' densities for the renew at propagation
        distance s (species2d%renew(s), simulation_class.f03:497-501)"""
        for sp in self.spe:
            if getattr(sp, "profile", None) is not None:
                sp.set_profile(sp.profile, s)

    def step3d_end(self):
        self._profiles_at((self.tstep + 1) * self.grid.dt)
        L.check(self.lib.qpad_step3d_end(self.ctx))
        self.tstep += 1

    def step3d(self):
        it = C.c_int32(0)
        self._profiles_at((self.tstep + 1) * self.grid.dt)
        L.check(self.lib.qpad_step3d(self.ctx, C.byref(it)))
        self.tstep += 1
        return it.value

    def run(self, nstep3d=None, diag=None):
        """simulation%run (simulation_class.f03:226-520).  With diag (a
        qpad_amd.diag.Diagnostics): on a fresh run the beam charge is
        deposited and the step-0 files written (:276-289); each 3D step
        writes its files after the beam push and before the species are
        renewed (:487-501)."""
        n = self.nstep3d if nstep3d is None else nstep3d
        total = 0
        if diag is None:
            for _ in range(n):
                total += self.step3d()
            return total
        if self.tstep == 0:
            L.check(self.lib.qpad_beam_qdp(self.ctx))
            diag.run(0)
        for _ in range(n):
            self.step3d_begin()
            for j in range(1, self.grid.nxi + 1):
                total += self.slice(j)
            L.check(self.lib.qpad_beam_push(self.ctx))
            self.tstep += 1
            diag.run(self.tstep)
            self._profiles_at(self.tstep * self.grid.dt)
            for sp in self.spe:
                sp.renew()
        return total

    def iteration_policy(self, adaptive=True, margin=0):
        """How the fused slice queues its predictor-corrector iterations
        (results do not depend on it; include/qpad_hip.h)."""
        L.check(self.lib.qpad_iteration_policy(self.ctx, int(bool(adaptive)), int(margin)))

    def launch_shape(self, dep_cells=None):
        """Get (or set) the deposit workgroup shape (include/qpad_hip.h)."""
        sh = L.LaunchShape()
        L.check(self.lib.qpad_get_launch_shape(self.ctx, C.byref(sh)))
        if dep_cells is not None:
            sh.dep_cells = int(dep_cells)
            L.check(self.lib.qpad_set_launch_shape(self.ctx, C.byref(sh)))
        return sh.dep_cells

    def deposit_form(self, mfma=None):
        """Get (or set) the current deposit's accumulation form: True = FP64
        matrix cores, False = wave reduce-scatter (include/qpad_hip.h)."""
        sh = L.LaunchShape()
        L.check(self.lib.qpad_get_launch_shape(self.ctx, C.byref(sh)))
        if mfma is not None:
            sh.dep_mfma = int(bool(mfma))
            L.check(self.lib.qpad_set_launch_shape(self.ctx, C.byref(sh)))
        return bool(sh.dep_mfma)

    def tune_launch(self, reps=5):
        """Autotune the launch shape on the current state; returns the mean
        ms per candidate dep_cells (2, 4, 8) and the choice."""
        ms = np.zeros(3)
        L.check(self.lib.qpad_tune_launch(self.ctx, int(reps), L.dptr(ms)))
        return dict(zip((2, 4, 8), ms.tolist())), self.launch_shape()

    def convergence_tester(self, op):
        """'record' or 'compare' on b_spe component 2 (simulation_class.f03:522-606)."""
        rel, ab = C.c_double(0.0), C.c_double(0.0)
        L.check(self.lib.qpad_convergence_tester(self.ctx, 0 if op == "record" else 1,
                                                 C.byref(rel), C.byref(ab)))
        return rel.value, ab.value

    def step_iterations(self):
        """(iterations summed over the last 3D step's slices, slices of that
        step that ended at iter_max)"""
        t, m = C.c_int32(0), C.c_int32(0)
        L.check(self.lib.qpad_step_iterations(self.ctx, C.byref(t), C.byref(m)))
        return t.value, m.value

    def kernel_time(self, which):
        ms, n = C.c_double(0.0), C.c_int64(0)
        L.check(self.lib.qpad_kernel_time(self.ctx, int(which), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def kernel_time_reset(self, enable=True):
        L.check(self.lib.qpad_kernel_time_reset(self.ctx, int(enable)))

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.qpad_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Several independent Simulations stepped as one batch (parameter
    scans): their slices run side by side on the device, each on its own
    stream (include/qpad_hip.h, qpad_batch_*).  Results equal each
    simulation's own step3d."""

    MODES = {"eager": 0, "graph": 1, "eager_single": 2}

    def __init__(self, sims, graph=False):
        self.sims = list(sims)
        self.lib = self.sims[0].lib
        arr = (C.c_void_p * len(self.sims))(*[s_.ctx.value for s_ in self.sims])
        h = C.c_void_p()
        L.check(self.lib.qpad_batch_create(arr, len(self.sims), C.byref(h)))
        self.h = h
        self.mode(graph)

    def mode(self, graph):
        """graph: True/False, or one of 'eager' (host thread per simulation),
        'graph' (captured slice loops), 'eager_single' (one polling host
        thread)."""
        m = self.MODES[graph] if isinstance(graph, str) else int(bool(graph))
        L.check(self.lib.qpad_batch_mode(self.h, m))

    def tune(self, reps=5):
        """Autotune every instance's launch shape; returns the choices."""
        L.check(self.lib.qpad_batch_tune(self.h, int(reps)))
        return [s_.launch_shape() for s_ in self.sims]

    def step3d(self):
        its = np.zeros(len(self.sims), np.int32)
        L.check(self.lib.qpad_batch_step3d(self.h, L.iptr(its)))
        for s_ in self.sims:
            s_.tstep += 1
        return its

    def close(self):
        if getattr(self, "h", None):
            self.lib.qpad_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

