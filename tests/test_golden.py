"""The reference's own sample data for the Gaussian broadening
(convolve/spec.line -> convolve/output, and the .adia pair), copied
unchanged into tests/golden/convolve by tests/golden/make_convolve_fixtures.py.

The grid and constants are those of the reference's convolve/main.f90:
x_i = -0.1 + i * (0.4 - (-0.1)) / 2001, i = 0..2001, printed as x_i + 2.7807
(default-real literals: single precision), fwhm = 5e-3 eV.  The envelope is
printed with 8 decimals, so the tolerance is 1e-8 (5e-9 rounding)."""
import os

import numpy as np
import pytest

from oracle.oracle import sticks_envelope

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "convolve")
PAIRS = [("spec.line", "output"), ("spec.line.adia", "output.adia")]


def grid_and_sigma():
    xs, xe, es = (float(np.float32(v)) for v in (-0.1, 0.4, 2.7807))
    nt = 2000
    x = xs + np.arange(nt + 2) * ((xe - xs) / (nt + 1))
    fwhm = 5e-3
    return x, es, fwhm / (2.0 * np.sqrt(2.0 * np.log(2.0)))


@pytest.mark.parametrize("sticks,out", PAIRS)
def test_oracle_envelope_reproduces_reference_output(sticks, out):
    sl = np.loadtxt(os.path.join(HERE, sticks))
    ref = np.loadtxt(os.path.join(HERE, out))
    x, shift, sigma = grid_and_sigma()
    np.testing.assert_allclose(ref[:, 0], x + shift, atol=1e-8)
    env = sticks_envelope(sl[:, 0], sl[:, 1], x, sigma)
    assert np.max(np.abs(env - ref[:, 1])) < 1e-8


@pytest.mark.gpu
@pytest.mark.parametrize("sticks,out", PAIRS)
def test_device_convolve_reproduces_reference_output(sticks, out):
    """nv_convolve divides by the maximum (convolve.py:47); the reference's
    output is the undivided envelope, so both are compared after dividing by
    their maxima."""
    from nadvibs_amd.engine import Engine
    from nadvibs_amd.synth import workload
    sl = np.loadtxt(os.path.join(HERE, sticks))
    ref = np.loadtxt(os.path.join(HERE, out))
    x, _, sigma = grid_and_sigma()
    with Engine(workload("cfg1")) as eng:
        got = eng.convolve(sl[:, 0], sl[:, 1], x, sigma)
    assert np.max(np.abs(got - ref[:, 1] / ref[:, 1].max())) < 2e-8
