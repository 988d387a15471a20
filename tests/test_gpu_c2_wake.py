"""Why the C2 wake is ill-conditioned, shown on the library (the oracle's own
run of the same deck gives the same picture; the full-size parity tests
compare the two slice by slice there).

C2 drives a blowout (n_b/n_p = 8, sigma_r = 0.5) on a grid resolved 14 times
finer in r than in xi (dr = 12/4096, dxi = 0.04).  Ahead of the back of the
first bubble every slice converges in a few iterations, |B_perp| stays O(1)
and every electron keeps gamma - p_z near 1.  Where the bubble closes on the
axis (slice ~300) electrons reach gamma - p_z -> 0: the quasi-static weight
1/(gamma - p_z) of their push and deposit grows without bound, such a
particle crosses many radial cells per xi step, and the predictor-corrector
iteration (simulation_class.f03:370-399) stops converging: slices end at
iter_max.  From there on the step is ill-conditioned (a 1e-15 change of a
slice's input moves the oracle's own result by up to 1e-4, the full-size
parity tests): two runs that differ only in rounding follow different
paths behind the bubble.  In some, electrons with gamma - p_z < 0.05 gain
runaway transverse momenta (|p_perp| up to 3e5 in the oracle's run of this
step) and B_perp reaches 1e3 to 1e6; in others (this library's run of this
step) they do not.  The test asserts what every run shares.
"""
import numpy as np
import pytest

from qpad_amd import Simulation, configs

pytestmark = pytest.mark.gpu


def test_c2_wake_breaks_down_behind_the_first_bubble():
    sim = Simulation(configs.c2())
    g = sim.grid
    sim.step3d_begin()
    rows = []
    for j in range(1, 561):
        its = sim.slice(j)
        if j % 20 == 0 or 280 <= j <= 320 and j % 5 == 0:
            b = sim.fields["b_spe"].get()
            s = sim.spe[0].get()
            gz = s["gamma"] - s["p3"]
            rows.append((j, its, float(np.max(np.abs(b[0:2, :, 1:-1]))), float(gz.min()),
                         int(np.sum(gz < 0.05)), float(np.max(np.hypot(s["p1"], s["p2"])))))
    sim.close()
    print("slice iterations max|B_perp| min(gamma-p_z) n(gamma-p_z<0.05) max|p_perp|")
    for r in rows:
        print("%5d %3d %10.3e %10.3e %6d %10.3e" % r)
    ahead = [r for r in rows if r[0] <= 280]
    back = [r for r in rows if 300 <= r[0] <= 320]
    behind = [r for r in rows if r[0] >= 300]
    # ahead of the bubble's back: converged, O(1) fields, gamma - p_z > 0.5
    assert all(r[1] < g.iter_max for r in ahead)
    assert max(r[2] for r in ahead) < 2.0
    assert min(r[3] for r in ahead) > 0.5
    # at the back: iter_max in every sampled slice, gamma - p_z below 0.2
    assert all(r[1] == g.iter_max for r in back)
    assert max(r[3] for r in back) < 0.25
    # behind: electrons near gamma - p_z = 0 in most sampled slices (below
    # 0.1 in some), and a third or more of them at iter_max
    gzs = sorted(r[3] for r in behind)
    assert gzs[len(gzs) // 2] < 0.25 and gzs[0] < 0.1
    assert sum(r[1] == g.iter_max for r in behind) >= len(behind) / 3
