"""Franck-Condon start vector (generate_initialvec with shiftref != 0,
Source/secondary.f90:351-783).

The oracle's recursion is pinned by
  * the displaced oscillator (T = I, equal frequencies): coherent-state
    amplitudes (-sqrt(S))^n / sqrt(n!) with S = b d^2 / 2;
  * 2-D quadrature of <psi_m^(b)(T Q + d) | phi_n^(a)(Q)> with Duschinsky
    rotation, different frequencies, the ground and a singly excited initial
    state (the reference's recursion corresponds to initial coordinates
    Q' = T Q + d).
The GPU tests compare the library with the oracle.
"""
import math

import numpy as np
import pytest
from numpy.polynomial.hermite import hermval

from nadvibs_amd.synth import make_problem
from oracle.oracle import Oracle


def fc_problem(nfunc, a_cm, b_cm, d, theta=0.0, istate=None, nstates=1, weights=None):
    M = len(nfunc)
    p = make_problem(nstates, nfunc, 1, (1000, 1000), 0.0, 0.0, 0.0, seed=1)
    p.aomega = np.array(a_cm) / 219474.6313705
    p.bomega = np.array(b_cm) / 219474.6313705
    p.dvec = np.array(d, dtype=float)
    T = np.eye(M)
    if M >= 2 and theta:
        c, s = math.cos(theta), math.sin(theta)
        T[:2, :2] = [[c, -s], [s, c]]
    p.tmat = T.T.reshape(-1).copy()  # tmat((i-1)M + j) = T(j, i): column-major
    p.shiftref = 1
    p.initstate = list(istate) if istate is not None else [0] * M
    if weights is not None:
        p.statew = np.array(weights, dtype=float)
    return p, T


def quadrature_amplitudes(p, T, m=(0, 0)):
    """<psi_m^(b)(T Q + d) | phi_n^(a)(Q)> for all n of a 2-mode basis."""
    a, b, d = p.aomega, p.bomega, p.dvec
    L = 12.0 / np.sqrt(np.minimum(a, b).min())
    g = np.linspace(-L, L, 801)
    dq = g[1] - g[0]
    Q1, Q2 = np.meshgrid(g, g, indexing="ij")
    Qp = np.stack([T[0, 0] * Q1 + T[0, 1] * Q2 + d[0], T[1, 0] * Q1 + T[1, 1] * Q2 + d[1]])

    def ho(n, w, x):
        c = np.zeros(n + 1); c[n] = 1.0
        return (w / np.pi) ** 0.25 / math.sqrt(2.0 ** n * math.factorial(n)) * hermval(np.sqrt(w) * x, c) * \
            np.exp(-w * x * x / 2)

    psi = ho(m[0], b[0], Qp[0]) * ho(m[1], b[1], Qp[1])
    out = np.zeros(p.nfunc)
    for n1 in range(p.nfunc[0]):
        f1 = ho(n1, a[0], Q1)
        for n2 in range(p.nfunc[1]):
            out[n1, n2] = np.sum(psi * f1 * ho(n2, a[1], Q2)) * dq * dq
    return out.reshape(-1)


def _norm(v):
    return v / np.linalg.norm(v)


def test_displaced_oscillator_poisson():
    b = [1500.0, 700.0]
    d = [6.0, -9.0]
    p, _ = fc_problem([14, 16], b, b, d)
    q = Oracle(p).start_vector()
    # coherent state centred at Q = -d: alpha = -d sqrt(b/2), |alpha|^2 = S = b d^2 / 2
    alpha = -np.array(d) * np.sqrt(np.array(b) / 219474.6313705 / 2)
    n1, n2 = np.meshgrid(np.arange(14), np.arange(16), indexing="ij")
    amp = lambda al, n: al ** n / np.sqrt([math.factorial(int(k)) for k in n.reshape(-1)]).reshape(n.shape)
    ref = (amp(alpha[0], n1) * amp(alpha[1], n2)).reshape(-1)
    np.testing.assert_allclose(q, _norm(ref), atol=1e-12)


@pytest.mark.parametrize("m", [(0, 0), (1, 0), (2, 0)])
def test_duschinsky_quadrature(m):
    p, T = fc_problem([9, 8], [1400.0, 900.0], [1100.0, 1300.0], [4.0, -3.0], theta=0.35, istate=list(m))
    q = Oracle(p).start_vector()
    ref = quadrature_amplitudes(p, T, m)
    np.testing.assert_allclose(q, _norm(ref), atol=2e-10)


def test_weights_and_states():
    p, _ = fc_problem([6, 5], [1000.0, 800.0], [1000.0, 800.0], [3.0, 2.0], nstates=3, weights=[1.0, 0.5, 2.0])
    q = Oracle(p).start_vector().reshape(3, -1)
    w = np.array([1.0, 0.5, 2.0]) / np.linalg.norm([1.0, 0.5, 2.0])
    for s in range(3):
        np.testing.assert_allclose(q[s], w[s] * _norm(q[0]), atol=1e-14)
    assert abs(np.linalg.norm(q) - 1.0) < 1e-14


@pytest.mark.gpu
@pytest.mark.parametrize("m", [(0, 0, 0), (0, 2, 0)])
def test_gpu_fc_start_matches_oracle(m):
    from nadvibs_amd.engine import Engine
    p, _ = fc_problem([9, 8, 7], [1400.0, 900.0, 600.0], [1100.0, 1300.0, 650.0], [4.0, -3.0, 2.0], theta=0.35,
                      istate=list(m), nstates=2, weights=[1.0, 0.7])
    ref = Oracle(p).start_vector()
    with Engine(p) as eng:
        q = eng.start_vector()
    np.testing.assert_allclose(q, ref, atol=1e-13)
