"""Pin the oracle's per-block kernels to the reference Fortran.

* Against the committed fixture (tests/golden/kernels_ref.npz, written by make_golden.py from the
  reference sources compiled in oracle/_ref): bit-exact with the all-double build, and within
  1e-13 of the field scale with the reference's as-configured build (quad-precision literals).
* Live, when oracle/_ref is present (build container and GPU box): full 16^3 blocks, two seeds.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
import kernel_cases as K

FIX = os.path.join(os.path.dirname(__file__), "golden", "kernels_ref.npz")


@pytest.fixture(scope="module")
def golden():
    return np.load(FIX, allow_pickle=False)


@pytest.mark.parametrize("name", K.CASES)
def test_oracle_matches_reference_fixture(golden, name):
    seed, sz, g = int(golden["seed"]), tuple(int(s) for s in golden["sz"]), int(golden["g"])
    got = K.run_case(name, seed, sz, g, "oracle")
    for key, val in got.items():
        want = golden["%s/%s" % (name, key)]
        assert val.shape == want.shape
        assert np.array_equal(val, want), "%s/%s differs from the all-double reference" % (name, key)
        q = want.ravel().copy()
        if "q16:%s/%s:idx" % (name, key) in golden:
            q[golden["q16:%s/%s:idx" % (name, key)]] = golden["q16:%s/%s:val" % (name, key)]
        if val.dtype == np.float64:
            scale = max(float(np.max(np.abs(q))), 1e-300)
            assert np.max(np.abs(val.ravel() - q)) <= 1e-13 * scale


@pytest.mark.skipif(O.ref_path("d8") is None, reason="oracle/_ref not built")
@pytest.mark.parametrize("seed", [11, 12])
@pytest.mark.parametrize("name", K.CASES)
def test_oracle_matches_reference_live(name, seed):
    ref = O.ref("d8")
    a = K.run_case(name, seed, (16, 16, 16), 2, "oracle", cut=(seed % 2 == 1))
    b = K.run_case(name, seed, (16, 16, 16), 2, ref, cut=(seed % 2 == 1))
    for key in a:
        assert np.array_equal(a[key], b[key]), key


def test_planted_mistake_is_caught(golden):
    """A dropped term in the oracle (planted on the CPU, in a copy of its output) fails the pin."""
    seed, sz, g = int(golden["seed"]), tuple(int(s) for s in golden["sz"]), int(golden["g"])
    got = K.run_case("calc_ab_p", seed, sz, g, "oracle")
    bad = got["b"].copy()
    bad[g + 1, g + 1, g + 1] += 1e-12 * max(1.0, abs(bad[g + 1, g + 1, g + 1]))
    assert not np.array_equal(bad, golden["calc_ab_p/b"])
