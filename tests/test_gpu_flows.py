"""The flow past a sphere of C1 (8 x 4 x 4 blocks of 16^3, sphere D = 1, Re 100, dt 0.01, explicit
U, pressure to 1e-8, seeded initial noise) over 30 time steps with the mixed-precision pressure
solve and with the reference's FP64 Jacobi-BiCGSTAB: both solve every step to the same FP64
residual, so the observables the reference prints agree to far below their own size -- the drag
coefficient (PrintForceCID's pressure and viscous sums on the sphere) and max u_x, every 10 steps."""
import os
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_c1_drag_mixed_equals_fp64():
    mixed = bench.flow_run_leg("c1", 30, every=10)
    fp64 = bench.flow_run_leg("c1", 30, every=10, fp64=True)
    assert [h["step"] for h in mixed["history"]] == [10, 20, 30]
    for a, b in zip(mixed["history"], fp64["history"]):
        assert abs(a["cd"] - b["cd"]) <= 1e-5 * abs(b["cd"]), (a, b)
        assert abs(a["ux_max"] - b["ux_max"]) <= 1e-5 * abs(b["ux_max"]), (a, b)
        assert 1.0 < b["cd"] < 3.0 and 1.0 < b["ux_max"] < 1.5
    # the mixed solve needs far fewer pressure iterations per step
    assert mixed["pressure_iterations_mean"] * 5 < fp64["pressure_iterations_mean"]
