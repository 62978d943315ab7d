"""Pins orc_gridsf (op/CasaGridSF.cu:41-85) against closed forms and the
defining property of the prolate spheroidal function."""
import numpy as np
import pytest
from scipy.integrate import quad


def test_endpoint_is_first_coefficient(orc):
    # at nu = 1 the second branch reduces to P1[0] / Q1[0]
    assert orc.gridsf(1.0) == pytest.approx(4.028559e-3, abs=1e-15)


def test_centre_value(orc):
    # Schwab's approximation of the normalised function: psi(0) = 1 to 4e-7
    assert abs(orc.gridsf(0.0) - 1.0) < 4e-7


def test_branches_meet_at_075(orc):
    lo = orc.gridsf(np.nextafter(0.75, 0.0))
    hi = orc.gridsf(0.75)
    assert abs(lo - 0.0820334) < 1e-7
    assert abs(lo - hi) < 5e-8


def test_monotone_decreasing(orc):
    v = [orc.gridsf(x) for x in np.linspace(0, 1, 201)]
    assert all(a > b for a, b in zip(v, v[1:]))


def test_spheroidal_eigen_property(orc):
    """psi is the alpha=1, c=3*pi prolate spheroidal function: the finite
    Fourier transform of (1-x^2) psi(x) over [-1,1] is proportional to psi.
    A wrong coefficient anywhere in the table breaks this at >1e-4."""
    def F(x):
        return quad(lambda e: (1 - e * e) * orc.gridsf(abs(e)) * np.cos(3 * np.pi * e * x),
                    -1, 1, limit=200, epsabs=1e-13)[0]
    xs = np.linspace(0.0, 0.95, 9)
    r = np.array([F(x) / orc.gridsf(x) for x in xs])
    assert np.max(np.abs(r / r[0] - 1.0)) < 3e-6
