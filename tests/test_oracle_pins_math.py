"""The oracle's field solves, deposits, pushers and injection against the
mathematics that fixes them (tests/oracle_pins.py states what fixes each),
and proof that each pin sees a plausible mistake: every mutant of
tests/oracle_mutants.py, an oracle copy with one planted error, must fail
the pins it names while the built oracle passes them all."""
import pytest

import oracle_mutants
import oracle_pins
import pyoracle as po


@pytest.mark.parametrize("pin", sorted(oracle_pins.PINS))
def test_oracle_meets_pin(pin):
    oracle_pins.PINS[pin](po.lib())


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    out = tmp_path_factory.mktemp("oracle_mutants")
    return {name: po.load_library(path) for name, path in oracle_mutants.build_all(str(out)).items()}


@pytest.mark.parametrize("name", sorted(oracle_mutants.MUTANTS))
def test_pins_see_planted_mistake(name, mutant_libs):
    L = mutant_libs[name]
    for pin in oracle_mutants.MUTANTS[name][2]:
        with pytest.raises(AssertionError):
            oracle_pins.PINS[pin](L)


def test_every_pin_has_a_mutant():
    caught = {p for m in oracle_mutants.MUTANTS.values() for p in m[2]}
    assert caught == set(oracle_pins.PINS)
