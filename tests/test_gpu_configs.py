"""Parity at the benchmark workloads' full sizes, through the C ABI.

  * 1.38M-row operator (3 states, basis [12,10,10,8,8,6], ordr 2): the first
    20 Lanczos coefficients in FP64 and with FP32 H*v against the oracle's
    FP64 run; the whole 2000-iteration job with partial re-orthogonalisation
    in both precisions: finite, orthogonal (a sample of the kept basis), and
    converged roots of the mixed run within 1e-3 cm^-1 of FP64
    (Source/top.f90:672-842, :844-1005).
  * 88.5M-row operator (4 states, 10 modes, ordr 3): H*v at 4096 random rows
    against the oracle's row restatement (FP64 and FP32; the fast-block form,
    the default there, and the kernels without it), the symmetry
    x.(H y) = y.(H x) over the whole vector, and the first three iterations of
    both precisions against the oracle's recurrence computed exactly on the
    rows the Krylov vectors reach (the start vector lives on row 0, and each
    H*v adds at most ordr quanta).
  * spectrum-loss gradient on the 1.38M-row operator, 500 iterations,
    sigma 15 cm^-1, 4096 points: against central differences of the device
    forward, and FP32 H*v against FP64 within 1e-5.
Tolerances: FP64 1e-13 relative for H*v and 1e-11 for the coefficients,
FP32 2e-6 for H*v and 5e-6 for the coefficients.
"""
import itertools
import sys

import numpy as np
import pytest

from nadvibs_amd.engine import Engine
from nadvibs_amd.inputs import AU2WAVE
from nadvibs_amd.synth import workload
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def cfg2():
    p = workload("cfg2")
    _say(f"workload 1.38M rows: N={p.n}")
    ref = Oracle(p).lanczos(20)
    return p, ref


@pytest.mark.parametrize("prec,tol", [("f64", 1e-11), ("f32", 5e-6)])
def test_cfg2_first_coefficients_match_oracle(cfg2, prec, tol):
    p, ref = cfg2
    with Engine(p) as eng:
        _say(f"cfg2 first 20 iterations, {prec}, kernel {eng.hv_kernel(prec)}")
        eng.lanczos_init(niter=20, reorthog=0, prec=prec)
        eng.run(1, 20)
        a, b, _ = eng.coefficients()
    da = np.max(np.abs(a - ref["alpha"]) / np.abs(ref["alpha"]))
    db = np.max(np.abs(b - ref["beta"]) / np.abs(ref["beta"]))
    _say(f"  alpha {da:.2e}  beta {db:.2e}")
    assert da <= tol and db <= tol


def _gram_sample(eng, n):
    idx = sorted(set(list(range(1, n + 1, max(1, n // 40))) + list(range(max(1, n - 9), n + 1))))
    Q = np.array([eng.lanczos_vector(j) for j in idx])
    G = Q @ Q.T
    return np.max(np.abs(G - np.eye(len(idx))))


def test_cfg2_full_job_both_precisions():
    """2000 iterations, reorthog 1, chkorthog 100, in FP64 and with FP32 H*v."""
    p = workload("cfg2")
    out = {}
    with Engine(p) as eng:
        for prec in ("f64", "f32"):
            _say(f"cfg2 job, 2000 iterations, reorthog 1, {prec}")
            eng.lanczos_init(niter=2000, reorthog=1, chkorthog=100, prec=prec, seed=2024)
            eng.run(1, 2000)
            a, b, _ = eng.coefficients()
            assert np.all(np.isfinite(a)) and np.all(np.isfinite(b))
            ortho = _gram_sample(eng, 2000)
            _say(f"  max |Q^T Q - I| over a sample of the basis: {ortho:.2e}")
            assert ortho <= (1e-6 if prec == "f64" else 1e-4)
            out[prec] = eng.ritz(a, b)[:3]
    ev64, it64, bj64 = out["f64"]
    ev32, it32, bj32 = out["f32"]
    conv = np.nonzero(bj64 < 1e-9)[0]
    assert conv.size > 100
    dE = max(abs(ev32[np.argmin(np.abs(ev32 - ev64[k]))] - ev64[k]) for k in conv) * AU2WAVE
    _say(f"  {conv.size} converged roots, max |dE| {dE:.2e} cm^-1")
    assert dE <= 1e-3


# ---- the 88.5M-row operator -------------------------------------------------
@pytest.fixture(scope="module")
def cfg3():
    p = workload("cfg3")
    _say(f"workload 88.5M rows: N={p.n}")
    return p, Oracle(p)


def _form(form, monkeypatch):
    """The default H*v kernel at this size (the fast-block form), or the
    kernels that serve without it (NV_HV_FB=0: FP32 shaped, FP64 two-pass)."""
    if form == "fast-block":
        monkeypatch.delenv("NV_HV_FB", raising=False)
    else:
        monkeypatch.setenv("NV_HV_FB", "0")


def _check_form(eng, form, prec):
    info = eng.hv_kernel(prec)
    if form == "fast-block":
        assert info.startswith("generic fast-block"), info
    else:
        assert info.startswith("generic two-pass" if prec == "f64" else "generic shaped"), info
    return info


@pytest.mark.parametrize("form", ["fast-block", "no-fast-block"])
@pytest.mark.parametrize("prec,tol", [("f64", 1e-13), ("f32", 2e-6)])
def test_cfg3_hv_rows_match_oracle(cfg3, prec, tol, form, monkeypatch):
    p, orc = cfg3
    _form(form, monkeypatch)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(p.n)
    rows = np.sort(rng.choice(orc.dimen, 4096, replace=False))
    rows[:4] = [0, 1, orc.dimen - 2, orc.dimen - 1]  # the corners of the basis
    ref = orc.hv_rows(x, rows)
    with Engine(p) as eng:
        _say(f"cfg3 H*v {prec}: kernel {_check_form(eng, form, prec)}")
        y = eng.hv(x, prec)
    got = y.reshape(p.nstates, -1)[:, rows].T
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    _say(f"  relative error over 4096 rows: {err:.2e}")
    assert err <= tol


def test_cfg3_hv_is_symmetric(cfg3):
    """x.(H y) = y.(H x) in FP64 over all 88.5M elements (the inter-state
    blocks carry no all-even couplings, so the operator is symmetric)."""
    p, _ = cfg3
    rng = np.random.default_rng(9)
    x, y = rng.standard_normal(p.n), rng.standard_normal(p.n)
    with Engine(p) as eng:
        hx, hy = eng.hv(x, "f64"), eng.hv(y, "f64")
    lhs, rhs = float(x @ hy), float(y @ hx)
    scale = np.linalg.norm(x) * np.linalg.norm(hy)
    _say(f"cfg3 symmetry: |x.Hy - y.Hx| / (|x||Hy|) = {abs(lhs - rhs) / scale:.2e}")
    assert abs(lhs - rhs) <= 1e-13 * scale


def _rows_within(nfunc, qmax):
    """Row indices of the basis functions with at most qmax quanta in total."""
    nfunc = list(nfunc)
    M = len(nfunc)
    shft = [int(np.prod(nfunc[k + 1:])) for k in range(M)]
    out = []

    def rec(k, left, r):
        if k == M:
            out.append(r)
            return
        for n in range(min(nfunc[k] - 1, left) + 1):
            rec(k + 1, left - n, r + n * shft[k])

    rec(0, qmax, 0)
    return np.array(sorted(out), dtype=np.int64)


@pytest.mark.parametrize("form", ["fast-block", "no-fast-block"])
def test_cfg3_first_iterations_match_exact_recurrence(cfg3, form, monkeypatch):
    """alpha(1:3), beta(0:2): the oracle's recurrence evaluated on the rows the
    Krylov vectors reach (q_k lives within 3(k-1) quanta of row 0)."""
    p, orc = cfg3
    _form(form, monkeypatch)
    ns, D = p.nstates, orc.dimen
    q = orc.start_vector()
    q_prev = np.zeros_like(q)
    alpha, beta = [], [float(np.linalg.norm(q))]
    for k in range(1, 4):
        rows = _rows_within(p.nfunc, 3 * k)
        hq = orc.hv_rows(q, rows)                      # (rows, states)
        qr = q.reshape(ns, D)[:, rows].T
        a = float(np.sum(qr * hq))
        r = hq - a * qr - beta[-1] * q_prev.reshape(ns, D)[:, rows].T if k > 1 else hq - a * qr
        b = float(np.linalg.norm(r))
        alpha.append(a)
        beta.append(b)
        q_prev = q
        q = np.zeros_like(q)
        q.reshape(ns, D)[:, rows] = (r / b).T
    alpha, beta = np.array(alpha), np.array(beta)
    with Engine(p) as eng:
        for prec, tol in (("f64", 1e-11), ("f32", 5e-6)):
            eng.lanczos_init(niter=3, reorthog=0, prec=prec)
            _say(f"cfg3 first 3 iterations, {prec}, kernel {_check_form(eng, form, prec)}")
            eng.run(1, 3)
            a, b, _ = eng.coefficients()
            da = np.max(np.abs(a - alpha) / np.abs(alpha))
            db = np.max(np.abs(b - beta) / np.abs(beta))
            _say(f"  alpha {da:.2e}  beta {db:.2e}")
            assert da <= tol and db <= tol


# ---- spectrum-loss gradient on the 1.38M-row operator --------------------------
def test_cfg4_gradient_central_differences_and_precisions():
    """500 iterations with partial re-orthogonalisation, sigma 15 cm^-1 on
    4096 points over 0-6000 cm^-1 above the lowest root; target = the
    spectrum of the couplings scaled by 1 + N(0, 0.05).  FP32 H*v against
    FP64 over that window; central differences over 0-1500 cm^-1."""
    import torch

    from nadvibs_amd.adjoint import spectrum_of_parameters
    p = workload("cfg2")
    niter, sigma = 500, 15.0
    rng = np.random.default_rng(1004)
    w0 = torch.tensor(p.aomega, dtype=torch.float64)
    c0 = torch.tensor(p.concoef, dtype=torch.float64)
    pt0 = torch.tensor(p.potterms, dtype=torch.float64)
    grid = torch.linspace(0.0, 6000.0, 4096, dtype=torch.float64)
    with torch.no_grad():
        target = spectrum_of_parameters(p, w0, c0, pt0 * (1 + 0.05 * torch.from_numpy(
            rng.standard_normal(pt0.shape))), niter, grid, sigma, "lowest", reorthog=1)

    def loss(w, c, pt, prec="f64"):
        env = spectrum_of_parameters(p, w, c, pt, niter, grid, sigma, "lowest", prec=prec, reorthog=1)
        return torch.linalg.vector_norm(env - target)

    grads = {}
    for prec in ("f64", "f32"):
        w, c, pt = (t.clone().requires_grad_(True) for t in (w0, c0, pt0))
        L = loss(w, c, pt, prec)
        L.backward()
        grads[prec] = (w.grad.numpy().copy(), c.grad.numpy().copy(), pt.grad.numpy().copy())
        _say(f"cfg4 gradient {prec}: loss {float(L):.8g}")
    flat = {k: np.concatenate([g.reshape(-1) for g in v]) for k, v in grads.items()}
    rel = np.max(np.abs(flat["f32"] - flat["f64"])) / np.max(np.abs(flat["f64"]))
    _say(f"  FP32 vs FP64 gradient: {rel:.2e} of the largest component")
    assert rel <= 1e-5
    # Central differences of the FP64 forward, on the window whose roots have converged at 500
    # iterations (0-1500 cm^-1): the projected gradient is the derivative of the Krylov-space
    # spectrum, which equals that of the computed loss once the window's roots have converged.
    grid = torch.linspace(0.0, 1500.0, 4096, dtype=torch.float64)
    with torch.no_grad():
        target = spectrum_of_parameters(p, w0, c0, pt0 * (1 + 0.05 * torch.from_numpy(
            np.random.default_rng(1004).standard_normal(pt0.shape))), niter, grid, sigma, "lowest", reorthog=1)
    w, c, pt = (t.clone().requires_grad_(True) for t in (w0, c0, pt0))
    L = loss(w, c, pt)
    L.backward()
    gw, gpt = w.grad.numpy().copy(), pt.grad.numpy().copy()
    _say(f"cfg4 low window: loss {float(L):.8g}")
    picks = [("w", (int(np.argmax(np.abs(gw))),))]
    # among the coefficients present (a zero slot of an inter-state all-even monomial, perturbed,
    # makes the reference's lower-only operator non-symmetric)
    order = np.argsort(-np.abs(gpt).reshape(-1) * (pt0.numpy().reshape(-1) != 0))[:5]
    picks += [("pt", tuple(int(i) for i in np.unravel_index(j, gpt.shape))) for j in order]
    with torch.no_grad():
        for name, idx in picks:
            base = {"w": w0, "pt": pt0}[name]
            h = 1e-5 * abs(float(base[idx]))
            up, dn = base.clone(), base.clone()
            up[idx] += h
            dn[idx] -= h
            if name == "w":
                fd = (float(loss(up, c0, pt0)) - float(loss(dn, c0, pt0))) / (2 * h)
                got = float(gw[idx])
            else:
                fd = (float(loss(w0, c0, up)) - float(loss(w0, c0, dn))) / (2 * h)
                got = float(gpt[idx])
            _say(f"  d loss / d {name}{idx}: adjoint {got:.8e}, central difference {fd:.8e}")
            assert abs(got - fd) <= 1e-4 * abs(fd), (name, idx, got, fd)
