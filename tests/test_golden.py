"""The oracle against the committed fixtures written from the reference's
own compiled CPU classes (tests/golden/make_golden.py).  These run where the
reference sources are absent, so the oracle stays pinned everywhere."""
import os

import numpy as np
import pytest

from tests.golden import inputs

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_vectors.npz")


@pytest.fixture(scope="module")
def gold():
    with np.load(FIX, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def close(a, b, tol=1e-14):
    return np.max(np.abs(np.asarray(a) - b)) <= tol * np.max(np.abs(b))


def test_pair_spread_and_force(orc, gold):
    d = inputs.pair_case()
    rho = orc.spread(d["dims"], d["sigma"], d["rcut"], d["xyzq"], d["box"])
    assert close(rho, gold["pair_rho"])
    f = orc.interpolate(d["dims"], d["sigma"], d["rcut"], d["xyzq"], d["box"], d["phi"])
    assert close(f, gold["pair_force"])


def test_normalized_spread(orc, gold):
    d = inputs.rand_case()
    rho = orc.spread(d["dims"], d["sigma"], d["rcut"], d["xyzq"], d["box"], normalize=True)
    assert close(rho, gold["rand_rho"])


def test_greens(orc, gold):
    d = inputs.greens_case()
    assert close(orc.greens_mul(0, d["dims"], d["box"], d["sigma"], d["spec"]), gold["greens_gse"])
    assert close(orc.greens_mul(1, d["dims"], d["box"], d["sigma"], d["spec"]), gold["greens_g2"])


def test_convolutions(orc, gold):
    d = inputs.conv_case()
    assert close(orc.conv_run_x(d["dims"], d["kx"], d["grid"]), gold["conv_x"])
    d = inputs.ongrid_case()
    assert close(orc.spread_on_grid(d["dims"], d["sigma"], d["rcut"], d["box"], d["grid"]), gold["ongrid"])
