"""CPU: the committed fixtures of tests/golden.  oracle_small.npz (written by
make_oracle_fixtures.py) freezes the oracle's outputs on small seeded
inputs; the oracle must still reproduce them: bit for bit where the oracle
is sequential (TOD, SkyMap, focal plane), to 1e-12 relative where OpenMP
atomics reorder FP64 sums (the sky adjoint)."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "oracle_small.npz")


@pytest.fixture(scope="module")
def fx():
    return dict(np.load(FIX))


def test_crosspolar_fixture_inputs(oracle, fx):
    """The cross-polar case's beams are what inputs.crosspolar_fields and the
    oracle's PolBeam give, and they exercise the A2/B2 terms."""
    import make_oracle_fixtures as m
    x = m.xconvolver_case(oracle)
    assert np.array_equal(x["x_beams"], fx["x_beams"])
    b = fx["x_beams"]
    assert np.abs(b[2]).max() >= 0.1 * np.abs(b[1]).max()
    assert np.abs(b[6]).max() >= 0.1 * np.abs(b[5]).max()


def test_oracle_reproduces_its_fixtures(oracle, fx):
    import make_oracle_fixtures as m
    out = m.outputs(oracle, fx)
    for k in ("c_tod_a", "c_tod_b", "c_tod_a_interp", "c_tod_b_interp", "x_tod_a", "x_tod_b",
              "m_AtA", "m_AtD", "m_I", "m_Q", "m_U", "m_hits", "f_xdet", "f_ydet", "f_rotdet"):
        assert np.array_equal(out[k], fx[k]), k
    for k in ("c_gI", "c_gQ", "c_gU", "x_gI", "x_gQ", "x_gU", "x_gbeam"):
        assert np.allclose(out[k], fx[k], rtol=1e-12, atol=1e-12 * np.abs(fx[k]).max()), k
