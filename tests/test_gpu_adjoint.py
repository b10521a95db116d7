"""GPU: the library's reverse sweep, Krylov-projected gradient and spectrum
adjoint against the oracle (oracle/adjoint.py, itself pinned by complex-step
derivatives and central differences), checkpoint recomputation against the
kept basis, mixed precision against FP64, and the torch binding end to end
against central differences of the device forward.

Tolerances: FP64 reverse sweep vs oracle 1e-9 of the largest gradient
component (short runs: the forward runs differ by rounding only);
checkpointed vs kept basis exactly equal (the recomputation is checked bit
for bit); Krylov gradient vs oracle on the same basis 1e-10; FP32 H*v vs FP64
1e-5 of the largest component; central differences 1e-5 (short runs) and
1e-6 (the Krylov gradient at 80 iterations).
"""
import numpy as np
import pytest
import torch

from nadvibs_amd.adjoint import broadened_spectrum, lanczos_coefficients, spectrum_of_parameters
from nadvibs_amd.engine import Engine
from nadvibs_amd.synth import make_problem
from oracle.adjoint import (krylov_gradient, lanczos_adjoint, lanczos_basis, parameter_matrices, spectrum,
                            spectrum_adjoint)
from oracle.dense import dense_hamiltonian

pytestmark = pytest.mark.gpu


def _prob(ns=2, nf=(5, 4, 4), ordr=2, offe=0.0, seed=31, quirks=3):
    return make_problem(ns, list(nf), ordr, (600, 2800), 250.0, 60.0, 40.0, cubic=15.0 if ordr > 2 else 0.0,
                        seed=seed, quirks=quirks, offdiag_even=offe)


def _flat(g, p):
    """Gradient dict in the oracle's parameter order (invalid POTterms slots dropped)."""
    from math import comb
    out = list(g["omega"]) + list(g["concoef"])
    for idx in np.ndindex(*p.potterms.shape):
        if idx[2] < comb(p.nmodes + idx[1], idx[1] + 1):
            out.append(g["potterms"][idx])
    return np.array(out)


def _gpu_grad(p, n, abar, bbar, prec="f64", checkpoint=0, generic=False):
    with Engine(p) as eng:
        if generic:
            eng.select_hv_kernel(True)
        eng.lanczos_init(n, prec=prec, savevecs=checkpoint == 0, checkpoint=checkpoint)
        eng.run(1, n)
        a, b, _ = eng.coefficients()
        start = eng.start_vector()
        g = eng.lanczos_adjoint(abar, bbar, prec)
    return a, b, start, g


@pytest.mark.parametrize("case", [dict(), dict(ns=3, nf=(3, 4, 3), ordr=3, offe=20.0),
                                  dict(ns=3, nf=(3, 4, 3), ordr=3, offe=20.0, quirks=0)],
                         ids=["ns2-ordr2", "ns3-ordr3-q2", "ns3-exact"])
def test_adjoint_matches_oracle(case):
    p = _prob(**case)
    n = 14
    rng = np.random.default_rng(7)
    abar, bbar = rng.standard_normal(n), rng.standard_normal(n + 1)
    a, b, start, g = _gpu_grad(p, n, abar, bbar)
    H = dense_hamiltonian(p)
    a_o, b_o, Q = lanczos_basis(H, start, n)
    np.testing.assert_allclose(a, a_o, rtol=1e-11, atol=1e-15)
    mats = [m for m in parameter_matrices(p) if m is not None]
    ref = lanczos_adjoint(H, Q, a_o, b_o, abar, bbar, mats)
    got = _flat(g, p)
    assert got.size == ref.size
    assert np.max(np.abs(got - ref)) <= 1e-9 * np.max(np.abs(ref))


def test_adjoint_checkpoints_equal_kept_basis():
    p = _prob(ns=3, nf=(8, 7, 6, 5), seed=33)
    n = 40
    rng = np.random.default_rng(8)
    abar, bbar = rng.standard_normal(n), rng.standard_normal(n + 1)
    _, _, _, g_keep = _gpu_grad(p, n, abar, bbar)
    for K in (7, 40, 64):
        _, _, _, g_cp = _gpu_grad(p, n, abar, bbar, checkpoint=K)
        for key in g_keep:
            np.testing.assert_array_equal(g_cp[key], g_keep[key])


@pytest.mark.parametrize("origin", ["lowest", "fixed"])
def test_spectrum_and_adjoint_match_oracle(origin):
    p = _prob()
    H = dense_hamiltonian(p)
    v0 = np.zeros(p.n)
    v0[np.arange(p.nstates) * p.dimen] = 1.0
    a, b, _ = lanczos_basis(H, v0, 40)
    grid = np.linspace(-300.0, 9000.0, 1000)
    e0 = a.min() - 0.002
    gbar = np.random.default_rng(10).standard_normal(grid.size)
    with Engine(p) as eng:
        env = eng.spectrum(a, b, grid, 30.0, origin, e0)
        ab, bb = eng.spectrum_adjoint(a, b, grid, 30.0, gbar, origin, e0)
    ref = spectrum(a, b, grid, 30.0, origin, e0)
    assert np.max(np.abs(env - ref)) <= 1e-10 * np.max(np.abs(ref))
    ra, rb = spectrum_adjoint(a, b, grid, 30.0, gbar, origin, e0)
    scale = max(np.max(np.abs(ra)), np.max(np.abs(rb)))
    assert np.max(np.abs(ab - ra)) <= 1e-8 * scale
    assert np.max(np.abs(bb - rb)) <= 1e-8 * scale


def test_torch_loss_gradient_matches_central_differences():
    """loss = |env - target|_2 through both autograd functions; a few
    coefficients, a concoef and every frequency by central differences of
    the device forward."""
    p = _prob(seed=35)
    n = 24
    grid = torch.linspace(-200.0, 8000.0, 600, dtype=torch.float64)
    sigma = 60.0
    rng = np.random.default_rng(11)
    target = torch.from_numpy(np.abs(rng.standard_normal(grid.numel())) * 0.05)
    omega = torch.tensor(p.aomega, dtype=torch.float64, requires_grad=True)
    concoef = torch.tensor(p.concoef, dtype=torch.float64, requires_grad=True)
    potterms = torch.tensor(p.potterms, dtype=torch.float64, requires_grad=True)
    dev = Engine(p)

    def loss_of(w, c, pt):
        a, b = lanczos_coefficients(p, w, c, pt, n)
        env = broadened_spectrum(a, b, grid, sigma, "fixed", float(p.concoef[0]), engine=dev)
        return torch.linalg.vector_norm(env - target)

    L = loss_of(omega, concoef, potterms)
    L.backward()
    checks = [("omega", k) for k in range(p.nmodes)] + [("concoef", 1)]
    nz = np.argwhere(np.abs(p.potterms) > 0)
    checks += [("potterms", tuple(nz[i])) for i in rng.choice(len(nz), 6, replace=False)]
    vals = {"omega": omega, "concoef": concoef, "potterms": potterms}
    for name, idx in checks:
        x = vals[name].detach()
        h = 1e-6 * max(abs(float(x[idx])), 1e-4)
        up, dn = [vals[k].detach().clone() for k in ("omega", "concoef", "potterms")], \
                 [vals[k].detach().clone() for k in ("omega", "concoef", "potterms")]
        j = ("omega", "concoef", "potterms").index(name)
        up[j][idx] += h
        dn[j][idx] -= h
        with torch.no_grad():
            fd = (float(loss_of(*up)) - float(loss_of(*dn))) / (2 * h)
        got = float(vals[name].grad[idx])
        assert abs(got - fd) <= 1e-5 * max(abs(fd), 1e-3 * abs(float(L))), (name, idx, got, fd)
    dev.close()


def _spec_loss_grad(p, n, prec, reorthog=0):
    grid = np.linspace(-100.0, 4000.0, 400)
    with Engine(p) as eng:
        eng.lanczos_init(n, reorthog=reorthog, prec=prec, savevecs=True)
        eng.run(1, n)
        a, b, _ = eng.coefficients()
        env = eng.spectrum(a, b, grid, 30.0, "fixed", 0.0)
        _, _, tb = eng.spectrum_adjoint(a, b, grid, 30.0, 2 * env, "fixed", 0.0, full=True)
        g = eng.krylov_gradient(tb)
        basis = np.array([eng.lanczos_vector(j) for j in range(1, n + 1)])
    return g, tb, basis


@pytest.mark.parametrize("reorthog", [0, 1])
def test_krylov_gradient_matches_oracle_on_the_same_basis(reorthog):
    p = _prob(ns=3, nf=(4, 4, 3, 3), seed=36, quirks=0)
    n = 60
    g, tb, basis = _spec_loss_grad(p, n, "f64", reorthog)
    mats = [m for m in parameter_matrices(p) if m is not None]
    ref = krylov_gradient(basis, tb, mats)
    got = _flat(g, p)
    assert np.max(np.abs(got - ref)) <= 1e-10 * np.max(np.abs(ref))


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_krylov_gradient_from_checkpoints_equals_kept_basis(prec):
    """Checkpoints every 7 iterations (60 = 8 full segments + 4): the segment
    pairs recomputed from the checkpoints give the kept-basis gradient (the
    recomputed vectors are the stored ones bit for bit; only the summation
    grouping differs).  Q2 operator with inter-state diagonal couplings, so
    dH/dtheta is not symmetric for those slots."""
    p = _prob(ns=3, nf=(5, 4, 3, 3), seed=38, quirks=3, offe=30.0)
    n = 60
    grid = np.linspace(-100.0, 4000.0, 400)
    out = []
    for kw in (dict(savevecs=True), dict(checkpoint=7)):
        with Engine(p) as eng:
            eng.lanczos_init(n, prec=prec, **kw)
            eng.run(1, n)
            a, b, _ = eng.coefficients()
            env = eng.spectrum(a, b, grid, 30.0, "fixed", 0.0)
            _, _, tb = eng.spectrum_adjoint(a, b, grid, 30.0, 2 * env, "fixed", 0.0, full=True)
            out.append((a, b, _flat(eng.krylov_gradient(tb), p)))
    (a1, b1, g1), (a2, b2, g2) = out
    assert np.array_equal(a1, a2) and np.array_equal(b1, b2)
    assert np.max(np.abs(g2 - g1)) <= 1e-12 * np.max(np.abs(g1))


def test_krylov_gradient_mixed_precision_within_1e5_of_fp64():
    p = _prob(ns=3, nf=(8, 7, 6, 5), seed=34, quirks=0)
    g64, _, _ = _spec_loss_grad(p, 120, "f64")
    g32, _, _ = _spec_loss_grad(p, 120, "f32")
    f64, f32 = _flat(g64, p), _flat(g32, p)
    assert np.max(np.abs(f32 - f64)) <= 1e-5 * np.max(np.abs(f64))


def test_spectrum_of_parameters_matches_central_differences():
    """The stable torch path at 80 iterations against central differences of
    the device forward (exact operator)."""
    p = _prob(ns=3, nf=(6, 5, 4, 4), seed=37, quirks=0)
    n = 80
    grid = torch.linspace(-100.0, 4000.0, 400, dtype=torch.float64)
    target = torch.zeros(grid.numel(), dtype=torch.float64)
    omega = torch.tensor(p.aomega, dtype=torch.float64, requires_grad=True)
    concoef = torch.tensor(p.concoef, dtype=torch.float64, requires_grad=True)
    potterms = torch.tensor(p.potterms, dtype=torch.float64, requires_grad=True)

    def loss_of(w, c, pt):
        env = spectrum_of_parameters(p, w, c, pt, n, grid, 30.0, "fixed", 0.0)
        return torch.sum((env - target) ** 2)

    L = loss_of(omega, concoef, potterms)
    L.backward()
    rng = np.random.default_rng(12)
    nz = np.argwhere(np.abs(p.potterms) > 0)
    checks = [("omega", (k,)) for k in range(p.nmodes)] + [("concoef", (0,)), ("concoef", (2,))]
    checks += [("potterms", tuple(nz[i])) for i in rng.choice(len(nz), 6, replace=False)]
    vals = {"omega": omega, "concoef": concoef, "potterms": potterms}
    names = ("omega", "concoef", "potterms")
    fds, gots = [], []
    for name, idx in checks:
        x = float(vals[name].detach()[idx])
        h = 1e-6 * max(abs(x), 1e-4)
        up = [vals[k].detach().clone() for k in names]
        dn = [vals[k].detach().clone() for k in names]
        j = names.index(name)
        up[j][idx] += h
        dn[j][idx] -= h
        with torch.no_grad():
            fds.append((float(loss_of(*up)) - float(loss_of(*dn))) / (2 * h))
        gots.append(float(vals[name].grad[idx]))
    fds, gots = np.array(fds), np.array(gots)
    assert np.max(np.abs(gots - fds)) <= 1e-6 * np.max(np.abs(fds)), (gots, fds)


def test_krylov_gradient_reports_a_basis_that_lost_orthogonality():
    """nv_krylov_gradient checks the basis it projects on: a run without
    re-orthogonalisation long enough to lose orthogonality returns
    NV_WARN_NONORTHOGONAL (a warning in the binding, NvError when strict),
    a re-orthogonalised run passes; checkpoints with re-orthogonalisation are
    refused (that run keeps every vector)."""
    from nadvibs_amd._lib import NvError
    p = _prob(ns=3, nf=(4, 4, 3, 3), seed=36, quirks=0)
    n = 60
    grid = np.linspace(-100.0, 4000.0, 400)
    for reorthog in (0, 1):
        with Engine(p) as eng:
            eng.lanczos_init(n, reorthog=reorthog, prec="f64", savevecs=True)
            eng.run(1, n)
            a, b, _ = eng.coefficients()
            env = eng.spectrum(a, b, grid, 30.0, "fixed", 0.0)
            _, _, tb = eng.spectrum_adjoint(a, b, grid, 30.0, 2 * env, "fixed", 0.0, full=True)
            if reorthog == 0:
                with pytest.warns(RuntimeWarning, match="lost orthogonality"):
                    eng.krylov_gradient(tb)
                with pytest.raises(NvError, match="lost orthogonality") as ex:
                    eng.krylov_gradient(tb, strict=True)
                assert ex.value.code == 6
            else:
                eng.krylov_gradient(tb, strict=True)
                assert eng.orthogonality_warning is None
        with Engine(p) as eng:
            with pytest.raises(NvError, match="checkpoints apply"):
                eng.lanczos_init(n, reorthog=1, prec="f64", checkpoint=10)
