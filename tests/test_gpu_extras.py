"""GPU: the PyTorch binding and the Gaussian broadening of convolve.py."""
import numpy as np
import pytest

from nadvibs_amd.engine import Engine
from nadvibs_amd.synth import workload
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def test_torch_hv_matches_oracle():
    import torch
    from nadvibs_amd import torch_ops
    p = workload("cfg1")
    x = np.random.default_rng(4).standard_normal(p.n)
    ref = Oracle(p).hv(x)
    with Engine(p) as eng:
        xt = torch.from_numpy(x).cuda()
        y = torch_ops.hv(eng, xt, "f64")
        torch.cuda.synchronize()
        got = y.cpu().numpy()
    assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 1e-13


def test_convolve_closed_form():
    """envelope(x) = sum_j y_j exp(-((x - x_j)/sigma)^2/2), divided by its
    maximum (convolve.py:28-48, sigma a standard deviation)."""
    rng = np.random.default_rng(2)
    xs = rng.uniform(0, 6000, 300)
    ys = rng.uniform(0, 1, 300)
    grid = np.linspace(0, 6000, 4096)
    env = np.exp(-0.5 * ((grid[:, None] - xs[None, :]) / 15.0) ** 2) @ ys
    env /= env.max()
    with Engine(workload("cfg1")) as eng:
        got = eng.convolve(xs, ys, grid, 15.0)
    np.testing.assert_allclose(got, env, rtol=1e-12, atol=1e-14)
