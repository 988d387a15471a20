"""qpinput.json reader (host logic, CPU)."""
import math

import pytest

from qpad_amd import QpadError, configs
from qpad_amd.input import beam_from_deck, grid_from_deck, species_from_deck, tri_gaussian_charge


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
    (lambda d: d["species"][0].update(push_type="standard"), "pusher"),
    (lambda d: d["species"][0].update(profile=["parabolic-channel", "uniform"]), "uniform"),
    (lambda d: d["simulation"].update(nlasers=1), "nlasers"),
    (lambda d: d["simulation"].update(field_boundary="periodic"), "boundary"),
    (lambda d: d["beam"][0].update(profile_type="standard"), "random"),
    (lambda d: d["beam"][0].update(push_type="boris"), "reduced"),
])
def test_unsupported_features_raise(edit, match):
    d = configs.small()
    edit(d)
    with pytest.raises(QpadError, match=match):
        g, _, z0 = grid_from_deck(d)
        species_from_deck(d["species"][0])
        beam_from_deck(d["beam"][0], z0)
