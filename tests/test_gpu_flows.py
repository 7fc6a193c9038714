"""The single-level flow past a sphere (8 x 4 x 4 blocks of 16^3, sphere D = 1, Re 100, dt 0.01, explicit
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


def test_c1_fields_against_oracle():
    """Two steps of the single-level flow from the uniform flow, pressure to 1e-10: the device's FP64 path and
    its mixed-precision path against the oracle's FP64 Solver::Update (C restatement, Jacobi-
    BiCGSTAB), field by field.  Velocities and face velocities agree to 1e-10 of their size; the
    pressure, which the solves fix only to their residual tolerance, to 1e-7, and its Laplacian lapp
    (which carries that error through two differences) to 1e-8."""
    d = bench.field_error_leg(steps=2)
    assert set(d["fields"]) == {"ux", "uy", "uz", "p", "lapp"} | {"v%d" % q for q in range(6)}
    for k, v in d["fields"].items():
        bound = {"p": 1e-7, "lapp": 1e-8}.get(k, 1e-10)
        for tag in ("fp64", "mixed"):
            assert v[tag]["rel"] <= bound, (k, tag, v[tag])
    assert max(d["mixed_pressure_iterations"]) * 5 < min(d["fp64_pressure_iterations"])
