"""HDF5 diagnostics of a run on the GPU, in QPAD's layout, against the
oracle's restatement of the same 3D step.

A small deck asks for field, species and beam diagnostics as QPAD's decks do
(input_file/*/qpinput.json 'diag' sections); Simulation.run writes them at
step 0 and after the step, after the beam push and before the species are
renewed (simulation_class.f03:276-289,487-501).  Every mode file of every
diagnostic is read back with the HDF5 library and compared with the
oracle's histories (fields, total and per-species charge, beam charge) and
particles (raw), all from the deck alone."""
import glob
import os

import numpy as np
import pytest

import h5read
import pyoracle as po
import twin
from conftest import rel_err
from qpad_amd import Simulation, configs, diag

pytestmark = pytest.mark.gpu

FIELDS = {"psi_cyl_m": ("Psi", "psi", po.PSI, 0), "ez_cyl_m": ("Ez", "ez", po.E, 2),
          "er_cyl_m": ("Er", "er", po.E, 0), "bphi_cyl_m": ("Bphi", "bphi", po.B, 1),
          "spec_br_cyl_m": ("Br_spec", "br", po.B_SPE, 0), "charge_cyl_m": ("Charge", "charge", po.Q_SPE, 0),
          "jz_cyl_m": ("Jz", "jz", po.CU, 2)}


def deck():
    d = configs.small(nr=32, nxi=24, max_mode=2, ppc=(2, 4), total_num=1 << 10, sigma_xi=0.1, offset_x=0.05,
                      ions=True, ion_ppc=(1, 2))
    d["field"] = {"diag": [{"name": list(FIELDS), "ndump": 1}]}
    for sp in d["species"]:
        sp["diag"] = [{"name": ["charge_cyl_m"], "ndump": 1}, {"name": ["raw"], "ndump": 1, "psample": 1}]
    d["beam"][0]["diag"] = [{"name": ["charge_cyl_m"], "ndump": 1}, {"name": ["raw"], "ndump": 1, "psample": 3}]
    return d


def modes(h, nm):
    """oracle history [slice][k][node] -> {mode folder: (nxi, nr) array}"""
    out = {}
    for k in range(nm):
        name = "Re0" if k == 0 else ("Re" if k % 2 else "Im") + str((k + 1) // 2)
        out[name] = h[:, k, 1:-1]
    return out


def check_mode_files(folder, dataname, want, it, tol=1e-10):
    """every mode file against the oracle, norm-wise over the diagnostic: the
    error of each mode array relative to the largest value of any mode (a
    mode far smaller than m = 0, such as the ions' Im1 charge, holds only the
    round-off of the free-running step at that scale)"""
    scale = max(float(np.max(np.abs(a))) for a in want.values())
    for sub, arr in want.items():
        path = os.path.join(folder, sub, f"{dataname}_{it:08d}.h5")
        with h5read.File(path) as f:
            got = f.dataset(dataname)
            assert f.attr("/", "ITER") == it
        assert got.shape == arr.shape, (path, got.shape, arr.shape)
        e = rel_err(got, arr, floor=max(scale, 1e-6))
        assert e <= tol, (path, e)


def test_run_diagnostics_match_oracle(tmp_path):
    d = deck()
    sim = Simulation(d)
    dg = diag.Diagnostics(sim, str(tmp_path))
    sim.run(1, diag=dg)
    nr, nxi, nm = sim.grid.nr, sim.grid.nxi, sim.nm
    sim.close()
    # the oracle: the same deck, one step, beam pushed, species not renewed
    o = twin.oracle_from_deck(d)
    o.step_begin()
    for j in range(1, nxi + 1):
        o.slice(j)
    o.L.ora_beam_push(o.h, 0)
    for name, (folder, dn, fid, comp) in FIELDS.items():
        h = o.hist(fid)[1:nxi + 1, comp]
        check_mode_files(os.path.join(tmp_path, "Fields", folder), dn, modes(h, nm), 1)
    for k in range(2):
        h = o.species_q_hist(k)[1:nxi + 1, 0]
        check_mode_files(os.path.join(tmp_path, f"Species{k + 1}", "Charge"), "charge", modes(h, nm), 1)
        with h5read.File(os.path.join(tmp_path, f"Species{k + 1}", "Raw", "raw_00000001.h5")) as f:
            x, p, q = f.dataset("x"), f.dataset("p"), f.dataset("q")
            assert f.attr("/", "tp") == len(q)
        s = o.species(k)
        assert len(q) == len(s["q"])
        for col, key in ((x[:, 0], "x1"), (x[:, 1], "x2"), (p[:, 2], "p3"), (q, "q")):
            assert rel_err(np.sort(col), np.sort(s[key])) <= 1e-10, key
    hq = o.hist(po.Q_BEAM)[1:nxi + 1, 0]
    check_mode_files(os.path.join(tmp_path, "Beam1", "Charge"), "charge", modes(hq, nm), 1)
    # step-0 files: the initial beam charge
    check_mode_files(os.path.join(tmp_path, "Beam1", "Charge"), "charge", modes(hq, nm), 0)
    with h5read.File(os.path.join(tmp_path, "Beam1", "Raw", "raw_00000001.h5")) as f:
        tp = f.attr("/", "tp")
        b3 = f.dataset("x3")
    nb = len(o.beam(0)["x3"])
    assert tp == nb // 3 and len(b3) == tp
    z0 = d["simulation"]["box"]["z"][0]
    assert np.min(b3) >= z0 and np.max(b3) <= z0 + nxi * sim.grid.dxi
    o.close()
    # one folder per mode array of every grid diagnostic
    assert sorted(os.path.basename(p) for p in glob.glob(os.path.join(tmp_path, "Fields", "Psi", "*"))) == \
        ["Im1", "Im2", "Re0", "Re1", "Re2"]


def test_neutral_diagnostics_match_oracle(tmp_path):
    """neutrals(1) charge_cyl_m, ion_cyl_m and raw (diagnostics_class.f03:
    324-390, 909-928) of an ionising gas beside a plasma, after one step"""
    d = configs.ionization(element=7, ion_max=5, plasma=True, n0=1e18, n_peak=8.0, max_mode=2, nr=32, nxi=48)
    d["neutrals"][0]["diag"] = [{"name": ["charge_cyl_m", "ion_cyl_m"], "ndump": 1},
                                {"name": ["raw"], "ndump": 1, "psample": 2}]
    sim = Simulation(d)
    dg = diag.Diagnostics(sim, str(tmp_path))
    sim.run(1, diag=dg)
    nxi, nm = sim.grid.nxi, sim.nm
    isp = sim.neut[0].id
    sim.close()
    o = twin.oracle_from_deck(d)
    o.step_begin()
    for j in range(1, nxi + 1):
        o.slice(j)
    h = o.neutral_rho_hist(isp)[1:nxi + 1, 0]
    assert np.max(np.abs(h)) > 0, "the gas must ionise"
    check_mode_files(os.path.join(tmp_path, "Neutral1", "Ion_charge"), "ion_charge", modes(h, nm), 1)
    hq = o.species_q_hist(isp)[1:nxi + 1, 0]
    check_mode_files(os.path.join(tmp_path, "Neutral1", "Charge"), "charge", modes(hq, nm), 1)
    with h5read.File(os.path.join(tmp_path, "Neutral1", "Raw", "raw_00000001.h5")) as f:
        x, q = f.dataset("x"), f.dataset("q")
        assert f.attr("/", "tp") == len(q)
    s = o.species(isp)
    assert len(q) == len(s["q"]) // 2 > 0
    # step 0: nothing ionised yet
    with h5read.File(os.path.join(tmp_path, "Neutral1", "Ion_charge", "Re0", "ion_charge_00000000.h5")) as f:
        assert np.all(f.dataset("ion_charge") == 0.0)
    o.close()
