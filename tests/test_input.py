"""qpinput.json reader (host logic, CPU)."""
import glob
import math
import os

import pytest

from qpad_amd import QpadError, configs
from qpad_amd._lib import BeamStdCfg
from qpad_amd.input import (FileBeam, beam_from_deck, grid_from_deck, load_deck, neutral_from_deck, place_file_beam,
                            species_from_deck, strip_comments, tri_gaussian_charge)

# QPAD's own example decks, where the reference tree is present (this
# container; the GPU runs use configs.tri_gaussian(), which restates one)
DECKS = "/root/reference/input_file"


def test_c2_deck_maps_to_grid():
    g, nstep, z0 = grid_from_deck(configs.c2())
    assert (g.nr, g.nxi, g.max_mode) == (4096, 1024, 2)
    assert g.dr == pytest.approx(12.0 / 4096)
    assert g.dxi == pytest.approx(0.04)
    # relax_fac default 1e-3 (dr / 0.02)^2 (sim_fields_class.f03:137) gives
    # alpha = relax_fac / dr^2 = 2.5 for any dr
    assert g.relax_fac / g.dr ** 2 == pytest.approx(2.5)
    assert nstep == 10 and z0 == pytest.approx(-1024 * 0.04 / 2)


def test_species_and_beam_sections():
    d = configs.c3()
    _, _, z0 = grid_from_deck(d)
    e = species_from_deck(d["species"][0])
    i = species_from_deck(d["species"][1])
    assert (e.ppc1, e.ppc2, e.num_theta) == (4, 1, 16)
    assert i.qm == 1.0 and i.mass == 1836.0 and i.neutralized == 0
    b = beam_from_deck(d["beam"][0], z0)
    assert b.total_num == 1 << 22
    assert b.center[0] == pytest.approx(0.05)          # offset 0.1 sigma_r
    assert b.center[2] == pytest.approx(5.0)           # 5 sigma_xi behind the box front
    assert b.total_charge == pytest.approx(8.0 * (2 * math.pi) ** 1.5 * 0.5 * 0.5 * 1.0)


def test_total_charge_definition():
    assert tri_gaussian_charge(4.0, (0.25, 0.25, 0.5)) == pytest.approx(1.96870, rel=1e-5)


@pytest.mark.parametrize("edit,match", [
    (lambda d: d["species"][0].update(push_type="boris"), "pusher"),
    (lambda d: d["species"][0].update(profile=["gaussian-channel", "uniform"]), "Invalid density profile"),
    (lambda d: d["species"][0].update(profile=["parabolic-channel", "uniform"]), "channel_n0"),
    (lambda d: d["simulation"].update(nlasers=1), "nlasers"),
    (lambda d: d["species"][0].update(smooth_order=2), "smooth_order"),
    (lambda d: d["simulation"].update(field_boundary="periodic"), "boundary"),
    (lambda d: d["beam"][0].update(profile_type="gaussian"), "Invalid beam profile type"),
    (lambda d: d["beam"][0].update(profile_type="standard", profile=["analytic"] * 3), "math_func"),
    (lambda d: d["beam"][0].update(profile_type="standard", profile=["gaussian", "cosine", "uniform"]),
     "Invalid beam density profile"),
    (lambda d: d["beam"][0].update(push_type="boris"), "reduced"),
])
def test_unsupported_features_raise(edit, match):
    d = configs.small()
    edit(d)
    with pytest.raises(QpadError, match=match):
        g, _, z0 = grid_from_deck(d)
        species_from_deck(d["species"][0])
        beam_from_deck(d["beam"][0], z0)


def test_comments_are_stripped_as_json_fortran_does():
    text = '{\n"a" : 1, !note\n"b" : "x!y" ! trailing\n}'
    assert strip_comments(text) == '{\n"a" : 1, \n"b" : "x!y" \n}'


def test_lattice_beam_section():
    d = configs.tri_gaussian()
    _, _, z0 = grid_from_deck(d)
    b = beam_from_deck(d["beam"][0], z0)
    assert isinstance(b, BeamStdCfg)
    assert list(b.ppc) == [2, 2, 2] and b.num_theta == 16 and b.quiet == 1
    assert list(b.profile) == [1, 1, 1]
    assert b.gauss_center[2] == pytest.approx(2.5) and b.range_lo[2] == 0.0 and b.range_hi[2] == 5.0
    # Twiss and offsets as QPAD sets them (fdist3d_std_class.f03:367-399)
    d["beam"][0].update(alpha=[0.5, -0.2], perp_offset_x=[-1.0, 0.1, 0.02])
    b = beam_from_deck(d["beam"][0], z0)
    assert b.twiss_beta[0] == pytest.approx(20000 * 0.25 / 5.0)
    assert b.offset_n[0] == 2 and b.offset_xi0[0] == pytest.approx(4.0) and b.offset_n[1] == 0


def _example_decks():
    return sorted(glob.glob(os.path.join(DECKS, "**", "*.json"), recursive=True))


@pytest.mark.skipif(not os.path.isdir(DECKS), reason="QPAD's example decks are not in this tree")
def test_reference_decks_load_or_are_refused_clearly():
    """Every example deck shipped with QPAD either maps onto the GPU path or
    is refused with the feature it needs (lasers, neutrals, analytic
    profiles, PGC pushers)."""
    ok, refused = [], {}
    for path in _example_decks():
        d = load_deck(path)
        try:
            g, _, z0 = grid_from_deck(d)
            for sp in d.get("species", [])[:d["simulation"].get("nspecies", 0)]:
                species_from_deck(sp)
            for bm in d.get("beam", [])[:d["simulation"].get("nbeams", 0)]:
                beam_from_deck(bm, z0)
            for nt in d.get("neutrals", [])[:d["simulation"].get("nneutrals", 0)]:
                neutral_from_deck(nt, d["simulation"].get("n0", 1.0e16))
            ok.append(os.path.relpath(path, DECKS))
        except QpadError as e:
            refused[os.path.relpath(path, DECKS)] = str(e)
    assert {"blowout_regime/qpinput_tri-gaussian.json", "blowout_regime/qpinput_sawtooth.json",
            "hosing/qpinput.json", "ionization/qpinput.json",
            "analytic_density_profiles/qpinput.json"} <= set(ok), (ok, refused)
    assert refused, "decks with lasers, neutrals or analytic profiles must be refused"
    for name, msg in refused.items():
        assert any(w in msg for w in ("nlasers", "nneutrals", "analytic", "pusher", "profile")), (name, msg)


@pytest.mark.skipif(not os.path.isdir(DECKS), reason="QPAD's example decks are not in this tree")
def test_analytic_config_restates_the_deck():
    real = load_deck(os.path.join(DECKS, "analytic_density_profiles", "qpinput.json"))
    mine = configs.analytic_example()
    for sec in ("beam", "species"):
        for a, b in zip(real[sec], mine[sec]):
            assert {k: v for k, v in a.items() if k != "diag"} == b
    for k, v in mine["simulation"].items():
        assert real["simulation"][k] == v, k


@pytest.mark.skipif(not os.path.isdir(DECKS), reason="QPAD's example decks are not in this tree")
def test_ionization_config_restates_the_deck():
    real = load_deck(os.path.join(DECKS, "ionization", "qpinput.json"))
    mine = configs.ionization_example()
    for sec in ("beam", "neutrals"):
        assert len(real[sec]) == len(mine[sec])
        for a, b in zip(real[sec], mine[sec]):
            assert {k: v for k, v in a.items() if k != "diag"} == b
    for k, v in mine["simulation"].items():
        assert real["simulation"][k] == v, k


def test_neutral_section():
    """neutrals(i) (sim_plasma_class.f03:148-176): element, ion_max, pusher,
    ppc and num_theta; n0 comes from the simulation section"""
    d = configs.ionization_example()
    c = neutral_from_deck(d["neutrals"][0], d["simulation"]["n0"])
    assert (c.ppc1, c.ppc2, c.num_theta, c.element, c.ion_max) == (8, 8, 32, 3, 3)
    assert c.push_type == 1 and c.qm == -1.0 and c.n0 == 1.0e17
    for edit, match in [(dict(element=4), "element 4"), (dict(ion_max=4), "ion_max 4"),
                        (dict(push_type="robust_pgc"), "pusher"), (dict(random_theta=True), "random_theta"),
                        (dict(profile=["parabolic-channel", "uniform"]), "uniform")]:
        nt = dict(d["neutrals"][0], **edit)
        with pytest.raises(QpadError, match=match):
            neutral_from_deck(nt, 1e17)


@pytest.mark.skipif(not os.path.isdir(DECKS), reason="QPAD's example decks are not in this tree")
def test_tri_gaussian_config_restates_the_deck():
    real = load_deck(os.path.join(DECKS, "blowout_regime", "qpinput_tri-gaussian.json"))
    mine = configs.tri_gaussian()
    for sec in ("beam", "species"):
        for a, b in zip(real[sec], mine[sec]):
            assert {k: v for k, v in a.items() if k != "diag"} == b
    for k, v in mine["simulation"].items():
        assert real["simulation"][k] == v, k


def test_file_beam_section():
    """beam(i) with profile_type "file" (init_fdist3d_file,
    fdist3d_file_class.f03:79-107): beam_center(3) made relative to the box
    start, conversion factors defaulting to 1."""
    bm = {"profile_type": "file", "filename": "beam.h5", "beam_center": [0.1, -0.2, 5.0],
          "file_center": [1.0, 2.0, 3.0], "q": -1.0, "m": 1.0, "charge_conv_fac": 2.5}
    fb = beam_from_deck(bm, z0=-2.0)
    assert isinstance(fb, FileBeam)
    assert fb.beam_center == (0.1, -0.2, 7.0) and fb.file_center == (1.0, 2.0, 3.0)
    assert (fb.length_conv, fb.charge_conv, fb.qm, fb.mass, fb.evol) == (1.0, 2.5, -1.0, 1.0, True)
    with pytest.raises(QpadError, match="spin"):
        beam_from_deck(dict(bm, has_spin=True), z0=0.0)


def test_file_beam_placement_by_hand():
    """inject_fdist3d_file's transform and box filter (fdist3d_file_class.
    f03:300-330) on particles whose images are worked out by hand: box
    r < 4 * 0.5 = 2, 0 <= xi < 10 * 0.25 = 2.5; centre (0.5, 0, 1), file
    centre (10, 20, 30), length factor 0.5, charge factor 3."""
    fb = FileBeam("f.h5", (0.5, 0.0, 1.0), (10.0, 20.0, 30.0), 0.5, 3.0, -1.0, 1.0, True)
    raw = {"x1": [10.0, 12.0, 10.0, 10.0, 10.0, 13.0, 9.0],
           "x2": [20.0, 20.0, 22.0, 20.0, 20.0, 20.0, 20.0],
           "x3": [30.0, 30.0, 30.0, 32.0, 28.0, 30.0, 28.0 - 1.0],
           "p1": [1, 2, 3, 4, 5, 6, 7], "p2": [0] * 7, "p3": [100.0] * 7, "q": [1, 2, 3, 4, 5, 6, 7]}
    # 0: at the file centre -> (0.5, 0, 1)
    # 1: x1 + 2 -> x1 = 0.5 + 1 = 1.5, r 1.5 kept
    # 2: x2 + 2 -> (0.5, 1), r = 1.118 kept
    # 3: x3 + 2 (toward the head) -> xi = 1 - 1 = 0, on the lower edge, kept
    # 4: x3 - 2 -> xi = 2 kept
    # 5: x1 + 3 -> x1 = 2.0, r = 2 on the upper edge, dropped
    # 6: x1 - 1 -> x1 = 0, x3 - 3 -> xi = 2.5 on the upper edge, dropped
    a = place_file_beam(raw, fb, nr=4, dr=0.5, nxi=10, dxi=0.25)
    assert a["x1"].tolist() == [0.5, 1.5, 0.5, 0.5, 0.5]
    assert a["x2"].tolist() == [0.0, 0.0, 1.0, 0.0, 0.0]
    assert a["x3"].tolist() == [1.0, 1.0, 1.0, 0.0, 2.0]
    assert a["p1"].tolist() == [1, 2, 3, 4, 5]
    assert a["q"].tolist() == [3.0, 6.0, 9.0, 12.0, 15.0]


def test_file_beam_placement_matches_the_oracle(oracle_mod):
    import numpy as np
    rng = np.random.default_rng(11)
    n = 4000
    raw = {k: rng.normal(0.0, 1.0, n) for k in ("x1", "x2", "x3", "p1", "p2", "p3", "q")}
    fb = FileBeam("f.h5", (0.2, -0.1, 1.5), (0.1, 0.0, -0.3), 0.7, 1e-3, -1.0, 1.0, True)
    mine = place_file_beam(raw, fb, nr=32, dr=0.05, nxi=64, dxi=0.05)
    ref = oracle_mod.file_beam_place(raw, fb.beam_center, fb.file_center, 0.7, 1e-3, 32, 0.05, 64, 0.05)
    assert 0 < len(mine["x1"]) < n
    for k in ref:
        assert np.array_equal(mine[k], ref[k]), k


def test_pusher_types():
    """species(i).push_type (sim_plasma_class.f03:123-137): "standard",
    "robust", "standard_pgc" and "robust_pgc" map onto the library's
    pushers (param.f03:63-64); the key is required, as in QPAD"""
    from qpad_amd._lib import PUSH2_ROBUST, PUSH2_ROBUST_PGC, PUSH2_STD, PUSH2_STD_PGC
    d = configs.small(push_type="standard", ions=True, ion_push_type="robust")
    assert species_from_deck(d["species"][0]).push_type == PUSH2_STD
    assert species_from_deck(d["species"][1]).push_type == PUSH2_ROBUST
    d = configs.small(push_type="standard_pgc", ions=True, ion_push_type="robust_pgc")
    assert species_from_deck(d["species"][0]).push_type == PUSH2_STD_PGC == 4
    assert species_from_deck(d["species"][1]).push_type == PUSH2_ROBUST_PGC == 5
    sp = dict(d["species"][0])
    del sp["push_type"]
    with pytest.raises(QpadError, match="push_type"):
        species_from_deck(sp)
