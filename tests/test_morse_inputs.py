"""Derivatives of the generated force constants with respect to the Morse parameters
(inputs.morse_ifc_derivatives), which feed the forward-mode derivatives of the library, and
the oracle fixture of the derivative checks."""
import os
import tempfile

import numpy as np
import pytest

from oracle import run as R
from tdep_amd import inputs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _flat(lists):
    return np.concatenate([np.ravel(e[-1]) for per in lists for e in per])


@pytest.mark.parametrize("name", ["rocksalt", "fcc_al"])
def test_force_constant_derivatives_equal_central_differences(name):
    p0, labels, potential, _ = inputs.morse_parameters(name)
    cr, _, cutoff = inputs.make_crystal(name)
    for k in range(len(p0)):
        h = 1e-6 * abs(p0[k])
        pp, pm = p0.copy(), p0.copy()
        pp[k] += h
        pm[k] -= h
        a2, a3 = inputs.morse_ifcs(cr, potential(pp), cutoff)
        b2, b3 = inputs.morse_ifcs(cr, potential(pm), cutoff)
        d2, d3 = inputs.morse_ifc_derivatives(name, p0, k)
        for fa, fb, fd in ((a2, b2, d2), (a3, b3, d3)):
            fdv, an = (_flat(fa) - _flat(fb)) / (2 * h), _flat(fd)
            assert np.max(np.abs(fdv - an)) <= 1e-7 * np.max(np.abs(an)), labels[k]


@pytest.mark.parametrize("name", ["rocksalt", "fcc_al"])
def test_force_constants_are_homogeneous_in_the_well_depths(name):
    """phi is linear in D (jitter factors included), so sum over the pair classes of
    D dfc/dD = fc for both orders, self terms included; a wrong chain factor on D (the
    jitter, the class selection) breaks it."""
    p0, _, potential, _ = inputs.morse_parameters(name)
    cr, _, cutoff = inputs.make_crystal(name)
    f2, f3 = inputs.morse_ifcs(cr, potential(p0), cutoff)
    s2, s3 = 0.0, 0.0
    for k in range(0, len(p0), 3):
        d2, d3 = inputs.morse_ifc_derivatives(name, p0, k)
        s2 = s2 + p0[k] * _flat(d2)
        s3 = s3 + p0[k] * _flat(d3)
    np.testing.assert_allclose(s2, _flat(f2), rtol=0, atol=1e-13 * np.abs(_flat(f2)).max())
    np.testing.assert_allclose(s3, _flat(f3), rtol=0, atol=1e-13 * np.abs(_flat(f3)).max())


def test_oracle_morse_fixture_reproduces():
    """The committed finite differences of the oracle: the frequency derivative of the first
    parameter, recomputed at the fixture's smallest step."""
    g = np.load(os.path.join(GOLDEN, "oracle_morse_rocksalt_6x6x6.npz"))
    p0 = g["params"]
    h = float(g["steps"][-1])
    out = []
    for s in (1 + h, 1 - h):
        p = p0.copy()
        p[0] *= s
        with tempfile.TemporaryDirectory() as d:
            inputs.make_config("rocksalt", d, params=p)
            out.append(R.setup(d, tuple(g["grid"]))[5].om_irr)
    fd = (out[0] - out[1]) / (2 * h * p0[0])
    np.testing.assert_allclose(fd, g["domega"][0], rtol=0, atol=1e-4 * np.abs(fd).max())
