"""Adjoint of the Lanczos recurrence and of the broadened spectrum, restated
densely in numpy (TEST INFRASTRUCTURE ONLY: the checker of the library's
nv_lanczos_adjoint / nv_spectrum_adjoint; nothing in the product imports it).

Forward map (the recurrence of Source/top.f90:646-670 and :820-842, without
re-orthogonalisation):
    u_i = H q_i,  alpha_i = q_i . u_i,  r_i = u_i - alpha_i q_i - beta_{i-1} q_{i-1},
    beta_i = |r_i|,  q_{i+1} = r_i / beta_i.
Its reverse sweep, i = n .. 1, given the adjoints abar_i, bbar_i of the
outputs and qbar_{i+1} accumulated so far:
    rbar   = (qbar_{i+1} - (qbar_{i+1} . q_{i+1}) q_{i+1}) / beta_i + bbar_i q_{i+1}
    atot   = abar_i - rbar . q_i
    ubar   = rbar + atot q_i
    qbar_i += -alpha_i rbar + atot u_i + H^T ubar
    qbar_{i-1} += -beta_{i-1} rbar,   bbar_{i-1} += -rbar . q_{i-1}
    dL/dtheta += ubar . (dH/dtheta) q_i.
Pinned by complex-step derivatives of the recurrence (analytic in H: dots
without conjugation, sqrt of the sum of squares), which have no subtraction
error.

The reverse sweep is the exact derivative of the computed coefficients, but it
amplifies rounding along converged Ritz directions (measured: FP64 gradients
of a 5040-dim problem are wrong by orders of magnitude at n = 80).  Spectral
losses therefore use the Krylov-projected gradient
    dL/dtheta = sum_jk tbar_jk q_j . (dH/dtheta) q_k,  tbar = dL/dT,
which is stable for any n and agrees with central differences of the
Lanczos-computed loss once the spectrum in the window has converged.

Spectrum (convolve.py:28-48 on the Ritz sticks of Source/secondary.f90:252-256):
env(x) = sum_i a_i^2 G(x - (theta_i - theta_ref) c), a_i = first component of
eigenvector i of T, c = AU2WAVE, G(u) = exp(-u^2 / (2 sigma^2)).  Because
env(x) = [f_x(T)]_00 with f_x(l) = G(x - (l - theta_ref) c), its gradient
with respect to T is S (D o a a^T) S^T with D_ik the divided difference
sum_x gbar(x) (f_x(theta_i) - f_x(theta_k)) / (theta_i - theta_k)
(Daleckii-Krein), which stays finite for the near-degenerate (ghost) Ritz
values of a long run.  Pinned by central differences in alpha and beta.
"""
import numpy as np

from nadvibs_amd.inputs import AU2WAVE
from oracle.dense import dense_hamiltonian


def lanczos_basis(H, v0, niter):
    """Lanczos in the library's operation order; returns alpha(1:n),
    beta(0:n) and the basis q_1 .. q_{n+1} as rows (q_{n+1} = r_n / beta_n)."""
    n = H.shape[0]
    dt = np.result_type(H, v0)
    alpha = np.zeros(niter, dtype=dt)
    beta = np.zeros(niter + 1, dtype=dt)
    Q = np.zeros((niter + 1, n), dtype=dt)
    beta[0] = np.sqrt(np.sum(v0 * v0))
    r = v0.astype(dt)
    q = np.zeros(n, dtype=dt)
    for i in range(1, niter + 1):
        q_prev = q
        q = r * (1.0 / beta[i - 1])
        Q[i - 1] = q
        w = H @ q
        alpha[i - 1] = np.sum(q * w)
        w = w - alpha[i - 1] * q
        if i > 1:
            w = w - beta[i - 1] * q_prev
        beta[i] = np.sqrt(np.sum(w * w))
        r = w
    Q[niter] = r / beta[niter]
    return alpha, beta, Q


def lanczos_adjoint(H, Q, alpha, beta, abar, bbar, dHs):
    """Reverse sweep above.  abar[n], bbar[n+1] (beta(0:n)); dHs: list of
    dense dH/dtheta.  Returns dL/dtheta for each."""
    n = alpha.size
    N = H.shape[0]
    bb = np.array(bbar, dtype=np.float64).copy()
    qb_next = np.zeros(N)
    qb = np.zeros(N)
    qb_prev = np.zeros(N)
    g = np.zeros(len(dHs))
    for i in range(n, 0, -1):
        qi, qn = Q[i - 1], Q[i]
        qp = Q[i - 2] if i > 1 else np.zeros(N)
        rbar = (qb_next - (qb_next @ qn) * qn) / beta[i] + bb[i] * qn
        atot = abar[i - 1] - rbar @ qi
        ubar = rbar + atot * qi
        u = H @ qi
        qb = qb - alpha[i - 1] * rbar + atot * u + H.T @ ubar
        if i > 1:
            bb[i - 1] -= rbar @ qp
            qb_prev = qb_prev - beta[i - 1] * rbar
        for t, dH in enumerate(dHs):
            g[t] += ubar @ (dH @ qi)
        qb_next, qb, qb_prev = qb, qb_prev, np.zeros(N)
    return g


def _zero_problem(prob):
    import copy
    z = copy.deepcopy(prob)
    z.potterms = np.zeros_like(prob.potterms)
    z.concoef = np.zeros_like(prob.concoef)
    return z


def parameter_matrices(prob):
    """dH/dtheta for theta = (aomega[k], concoef[b], potterms[b, p, k]) in that
    order, from dense builds: complex step in the frequencies, differences
    of unit-coefficient builds for the (linear) coefficients."""
    import copy
    out = []
    h = 1e-30
    for k in range(prob.nmodes):
        pc = copy.deepcopy(prob)
        w = np.asarray(prob.aomega, dtype=complex).copy()
        w[k] += 1j * h
        pc.aomega = w
        out.append(dense_hamiltonian(pc, dtype=complex).imag / h)
    z = _zero_problem(prob)
    H0 = dense_hamiltonian(z)
    for b in range(prob.nstblks):
        pz = copy.deepcopy(z)
        pz.concoef[b] = 1.0
        out.append(dense_hamiltonian(pz) - H0)
    for idx in np.ndindex(*prob.potterms.shape):
        if idx[2] >= _noterms(prob.nmodes, idx[1] + 1):
            out.append(None)
            continue
        pz = copy.deepcopy(z)
        pz.potterms[idx] = 1.0
        out.append(dense_hamiltonian(pz) - H0)
    return out


def _noterms(nm, p):
    from math import comb
    return comb(nm + p - 1, p)


def complex_step_lanczos(H, dH, v0, niter, h=1e-30):
    """d(alpha, beta)/dtheta along dH by a complex step through the recurrence."""
    a, b, _ = lanczos_basis(H.astype(complex) + 1j * h * dH, v0.astype(complex), niter)
    return a.imag / h, b.imag / h


# ---- spectrum -------------------------------------------------------------
def _eig(alpha, beta):
    n = alpha.size
    T = np.diag(alpha) + np.diag(beta[1:n], 1) + np.diag(beta[1:n], -1)
    return np.linalg.eigh(T)


def spectrum(alpha, beta, grid, sigma, origin="lowest", e0=0.0):
    """Envelope on grid (cm^-1) of the sticks ((theta_i - theta_ref) c, a_i^2);
    theta_ref = lowest Ritz value or e0 (hartree)."""
    th, S = _eig(np.asarray(alpha), np.asarray(beta))
    a = S[0]
    ref = th[0] if origin == "lowest" else e0
    x = (th - ref) * AU2WAVE
    G = np.exp(-0.5 * ((np.asarray(grid)[:, None] - x[None, :]) / sigma) ** 2)
    return G @ (a * a)


def spectrum_adjoint(alpha, beta, grid, sigma, gbar, origin="lowest", e0=0.0, full=False):
    """(abar, bbar) of L with dL/denv = gbar; bbar holds beta(0:n) (0 and n
    do not enter T).  full=True also returns dL/dT as a symmetric matrix."""
    alpha = np.asarray(alpha, dtype=np.float64)
    beta = np.asarray(beta, dtype=np.float64)
    n = alpha.size
    th, S = _eig(alpha, beta)
    a = S[0]
    ref = th[0] if origin == "lowest" else e0
    c = AU2WAVE
    u = (np.asarray(grid)[:, None] - (th - ref)[None, :] * c) / sigma
    G = np.exp(-0.5 * u * u)
    W = gbar @ G                                  # sum_x gbar f_x(theta_i)
    Wp = gbar @ (G * u) * (c / sigma)             # d/dtheta_i
    dth = th[:, None] - th[None, :]
    small = np.abs(dth) <= 1e-7 * sigma / c
    Dm = np.where(small, 0.5 * (Wp[:, None] + Wp[None, :]), (W[:, None] - W[None, :]) / np.where(small, 1.0, dth))
    M = Dm * np.outer(a, a)
    if origin == "lowest":
        # env depends on theta_ref = theta_1 through every stick position
        M[0, 0] += -np.sum(a * a * Wp)
    Tb = S @ M @ S.T
    abar = np.diag(Tb).copy()
    bbar = np.zeros(n + 1)
    bbar[1:n] = 2.0 * np.diag(Tb, 1)
    return (abar, bbar, Tb) if full else (abar, bbar)


def krylov_gradient(Q, tbar, dHs):
    """dL/dtheta = sum_jk tbar[j, k] q_j . (dH/dtheta) q_k over the basis rows
    Q (q_1..q_n): the derivative of v0^T f(H) v0 projected on the Krylov
    space, for a spectral loss with dL/dT = tbar."""
    n = tbar.shape[0]
    W = tbar @ Q[:n]
    return np.array([np.sum(W * (Q[:n] @ dH.T)) for dH in dHs])
