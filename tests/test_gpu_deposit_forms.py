"""The current deposit's two accumulation forms (include/qpad_hip.h,
qpad_launch_shape.dep_mfma): the wave reduce-scatter and the FP64 matrix-core
GEMM add the same per-particle terms (amjdeposit_robust_part2d /
amjdeposit_std_part2d, part2d_class.f03:877-960) in different orders.  Each
form matches the oracle slice by slice, the two agree to rounding over a
3D step, and each is run-to-run identical."""
import numpy as np
import pytest

import twin
from qpad_amd import Simulation, configs

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _deck(max_mode, push="robust", ion_push=None):
    ions = ion_push is not None
    ppc = (2, 4) if max_mode <= 2 else (2, 16)
    return configs.small(nr=48, nxi=40, max_mode=max_mode, ppc=ppc, total_num=1 << 12, ions=ions,
                         ion_ppc=(1, 10), offset_x=0.03, push_type=push, ion_push_type=ion_push)


@pytest.mark.parametrize("max_mode,push,ion_push,mfma", [
    (1, "robust", None, True),
    (2, "robust", "standard", True),
    (4, "standard", "robust", False),
    (4, "robust", "standard", True),
])
def test_deposit_form_matches_oracle(max_mode, push, ion_push, mfma):
    print("deposit form: max_mode", max_mode, push, ion_push, "mfma", mfma, flush=True)
    sim = Simulation(_deck(max_mode, push, ion_push))
    assert sim.deposit_form() == (max_mode >= 3)   # the default
    sim.deposit_form(mfma)
    assert sim.deposit_form() == mfma
    o = twin.make_twin(sim)
    sim.step3d_begin()
    o.step_begin()
    norm = twin.ErrorNorm()
    for j in range(1, sim.grid.nxi + 1):
        twin.sync_slice_inputs(sim, o, j)
        assert sim.slice(j) == o.slice(j), j
        twin.compare_fields(sim, o, norm, j)
        twin.compare_species(sim, o, norm, j)
    print({n: f"{e:.1e}" for n, e in sorted(norm.errors().items())})
    bad = norm.bad(TOL)
    assert not bad, f"beyond {TOL}: {bad}"
    sim.close()
    o.close()


def _step_record(deck, mfma):
    sim = Simulation(deck)
    sim.deposit_form(mfma)
    rec = twin.record_gpu_step(sim)
    sim.close()
    return rec


@pytest.mark.parametrize("max_mode", [2, 4])
def test_forms_agree_and_repeat(max_mode):
    deck = _deck(max_mode, "robust", "standard")
    rs = _step_record(deck, False)
    mf = _step_record(deck, True)
    mf2 = _step_record(deck, True)
    for k in mf:
        assert np.array_equal(mf[k], mf2[k]), f"matrix form not run-to-run identical: {k}"
        a, b = np.asarray(mf[k], dtype=np.float64), np.asarray(rs[k], dtype=np.float64)
        scale = max(float(np.max(np.abs(b))), 1e-300) if b.size else 1.0
        err = float(np.max(np.abs(a - b))) / scale if b.size else 0.0
        assert err <= 1e-11, (k, err)
