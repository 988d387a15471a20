"""The oracle's self-sensitivity probe used by the full-size parity tests
(CPU): with a zero perturbation the rerun of a slice is bit-identical (the
snapshot restores every input the slice reads), and in a weak wake a 1e-15
change of the input stays at round-off."""
import twin
from qpad_amd import configs


def test_zero_perturbation_reruns_bit_identically():
    deck = configs.small(nr=40, nxi=24, max_mode=2, ppc=(2, 4), total_num=1 << 12, ions=True, offset_x=0.03)
    o = twin.oracle_from_deck(deck)
    o.step_begin()
    for j in range(1, 15):
        o.slice(j)
    its, norm = twin.oracle_self_sensitivity(o, 15, 2, eps=0.0)
    assert its >= 1
    assert all(e == 0.0 for e in norm.errors().values()), norm.errors()
    its2, norm2 = twin.oracle_self_sensitivity(o, 16, 2, eps=1e-15)
    assert max(norm2.errors().values()) < 1e-11, norm2.errors()
    o.close()
