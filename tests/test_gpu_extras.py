"""GPU: the PyTorch binding, the Gaussian broadening of convolve.py and
the library's refusals of unsupported run options."""
import itertools

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


def test_bf16_with_reorthogonalisation_is_refused():
    """BF16 input rounding outgrows partial re-orthogonalisation on long runs,
    so the library refuses the combination (status 1, a message)."""
    with Engine(workload("cfg1")) as eng:
        with pytest.raises(RuntimeError, match="BF16"):
            eng.lanczos_init(niter=50, reorthog=1, prec="bf16")
        eng.lanczos_init(niter=50, reorthog=0, prec="bf16")
        eng.run(1, 50)
        a, b, _ = eng.coefficients()
        assert np.all(np.isfinite(a)) and np.all(np.isfinite(b))


def test_range_end_at_or_above_trigger_is_refused():
    """The block-scan range detection needs every index to be either small
    (|omega| < eta) or not; eta >= tau is refused after the defaults apply."""
    with Engine(workload("cfg1")) as eng:
        with pytest.raises(RuntimeError, match="pro_eta < pro_tau"):
            eng.lanczos_init(niter=50, reorthog=1, prec="f64", pro_tau=1e-13)  # below the default eta eps^(3/4) = 1.1e-12
        with pytest.raises(RuntimeError, match="pro_eta < pro_tau"):
            eng.lanczos_init(niter=50, reorthog=1, prec="f32", pro_tau=1e-12)  # below the mixed eta 2^-36
        with pytest.raises(RuntimeError, match="pro_eta < pro_tau"):
            eng.lanczos_init(niter=50, reorthog=1, prec="f64", pro_tau=1e-6, pro_eta=1e-6)
        eng.lanczos_init(niter=50, reorthog=1, prec="f64", pro_tau=1e-6, pro_eta=1e-9)


def _step_valid_brute_force(p):
    """Share of (row, step) pairs of the off-diagonal steps whose target row
    lies inside the basis, counted over every row."""
    from nadvibs_amd.inputs import monomials
    nf = np.array(p.nfunc)
    D = int(np.prod(nf))
    digits = np.array(np.unravel_index(np.arange(D), nf))      # (M, D)
    ns = p.nstates
    nb = ns * (ns + 1) // 2
    valid = total = 0
    seen = set()
    for order in range(1, p.ordr + 1):
        for k, tup in enumerate(monomials(p.nmodes, order)):
            if not np.any(np.abs(p.potterms[:nb, order - 1, k]) >= 1e-20):
                continue
            pows = np.bincount(np.array(tup), minlength=p.nmodes)
            modes = [m for m in range(p.nmodes) if pows[m] > 0]
            key = tuple(pows)
            if key in seen:
                continue
            seen.add(key)
            smax = []
            for m in modes:
                s_ = int(pows[m])
                while s_ > nf[m] - 1:
                    s_ -= 2       # the largest step of the power's parity inside the basis
                smax.append(s_)
            if min(smax) < 0:
                continue
            ranges = [range(-s, s + 1, 2) for s in smax]
            for d in itertools.product(*ranges):
                if all(v == 0 for v in d):
                    continue
                ok = np.ones(D, bool)
                for m, v in zip(modes, d):
                    n2 = digits[m] + v
                    ok &= (n2 >= 0) & (n2 < nf[m])
                valid += int(ok.sum())
                total += D
    return valid / total


@pytest.mark.parametrize("forced", ["0", str(1 << 62)])
def test_shaped_skip_form_is_reported_and_counts_steps(monkeypatch, forced):
    """nv_hv_kernel_info names the form of the shaped iteration kernel (the
    loads of out-of-basis steps pointed outside the vector, or not) and the
    share of in-basis steps that picks it; the share equals a brute-force
    count, and the first coefficients agree with the oracle in both forms."""
    from nadvibs_amd.synth import make_problem
    monkeypatch.setenv("NV_HV_KERNEL", "generic")
    monkeypatch.setenv("NV_HV_SKIP_MIN_BYTES", forced)
    p = make_problem(3, [5, 4, 3, 4], 3, (400, 3000), 250.0, 60.0, 40.0, cubic=10.0, seed=21)
    ref = Oracle(p).lanczos(12)
    with Engine(p) as eng:
        for prec, tol in (("f64", 1e-11), ("f32", 5e-6)):
            info = eng.hv_kernel(prec)
            print(f"{prec}: {info}", flush=True)
            assert ("skip form" if forced == "0" else "plain form") in info
            share = float(info.split("steps inside the basis")[1].strip(" ]"))
            assert abs(share - _step_valid_brute_force(p)) < 1e-6
            eng.lanczos_init(niter=12, reorthog=0, prec=prec)
            eng.run(1, 12)
            a, b, _ = eng.coefficients()
            assert np.max(np.abs(a - ref["alpha"]) / np.abs(ref["alpha"])) <= tol
            assert np.max(np.abs(b - ref["beta"]) / np.abs(ref["beta"])) <= tol
