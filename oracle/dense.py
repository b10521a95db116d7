"""Dense brute force of the vibronic Hamiltonian (TEST INFRASTRUCTURE ONLY).

Built independently of the stencil code and of the reference's closed-form
matrix elements: the truncated q^p matrix of a mode is taken from the p-th
power of the ladder-operator matrix q = (a + a^dagger)/sqrt(2) on a larger
basis, and every monomial becomes a Kronecker product.  The reference's
conventions are then imposed explicitly:
  * <m|q^3|m+1> at one third (Source/function.f90:225) and powers above 4
    zero (hoelem's case default) when the Q3 quirk is on;
  * the vibrationally diagonal part of an inter-state block only below the
    state diagonal (Source/top.f90:752-754) when the lower-diagonal quirk is
    on, and only for blocks with a constant above 1e-16 or an all-even
    monomial (Source/top.f90:528-553).
Monomials are enumerated with itertools in the nadvibs.in order (Source/top.f90:153-168).
"""
from itertools import combinations_with_replacement

import numpy as np

QUIRK_Q3 = 1
QUIRK_LOWER = 2


def qpow(n, p, quirks=0):
    """Truncated analytic <i|q^p|j>, i, j < n."""
    L = n + p + 2
    a = np.diag(np.sqrt(np.arange(1, L)), 1)
    q = (a + a.T) / np.sqrt(2.0)
    m = np.linalg.matrix_power(q, p)[:n, :n]
    if quirks & QUIRK_Q3:
        if p == 3:
            idx = np.arange(n - 1)
            m[idx, idx + 1] /= 3.0
            m[idx + 1, idx] /= 3.0
        if p > 4:
            m[:] = 0.0
    return m


def mode_tuples(nmodes, p):
    return [tuple(reversed(c)) for c in combinations_with_replacement(range(nmodes), p)]


def monomial_operator(nfunc, pows, quirks):
    out = np.ones((1, 1))
    for k, nf in enumerate(nfunc):
        f = qpow(nf, pows[k], quirks) if pows[k] else np.eye(nf)
        out = np.kron(out, f)
    return out


def dense_hamiltonian(prob, dtype=float):
    """H as an (N, N) array in the state-major layout x[s][n_1..n_M].
    dtype=complex accepts complex frequencies (for complex-step derivatives)."""
    ns, nm, nf = prob.nstates, prob.nmodes, list(prob.nfunc)
    D = int(np.prod(nf))
    H = np.zeros((ns * D, ns * D), dtype=dtype)
    # harmonic part
    # digits of every row (last mode fastest), without an nmodes-dimensional array
    r = np.arange(D)
    shft = [int(np.prod(nf[k + 1:])) for k in range(nm)]
    E = sum(((r // shft[k]) % nf[k] + 0.5) * prob.aomega[k] for k in range(nm))
    for s in range(ns):
        H[s * D:(s + 1) * D, s * D:(s + 1) * D] += np.diag(E)
    blocks = [(i, j) for i in range(ns) for j in range(i + 1)]
    active = {}
    for b, (i, j) in enumerate(blocks):
        c0 = prob.concoef[b]
        active[(i, j)] = abs(c0) > 1e-16
        blk = c0 * np.eye(D)
        for p in range(1, prob.ordr + 1):
            for k, tup in enumerate(mode_tuples(nm, p)):
                c = prob.potterms[b, p - 1, k]
                if not abs(c) >= prob.ztoler:
                    continue
                pows = np.bincount(np.array(tup), minlength=nm)
                used = pows[pows > 0]
                if np.all(used % 2 == 0):
                    active[(i, j)] = True
                kappa = c / np.prod(np.asarray(prob.aomega) ** (pows / 2.0))
                blk = blk + kappa * monomial_operator(nf, pows, prob.quirks)
        H[i * D:(i + 1) * D, j * D:(j + 1) * D] += blk
        if i != j:
            H[j * D:(j + 1) * D, i * D:(i + 1) * D] += blk
    for (i, j), act in active.items():
        if i == j:
            continue
        up = H[j * D:(j + 1) * D, i * D:(i + 1) * D]
        lo = H[i * D:(i + 1) * D, j * D:(j + 1) * D]
        if not act:
            np.fill_diagonal(up, 0.0)
            np.fill_diagonal(lo, 0.0)
        elif prob.quirks & QUIRK_LOWER:
            np.fill_diagonal(up, 0.0)
    return H


def dense_lanczos(H, v0, niter):
    """Textbook Lanczos in the reference's operation order (alpha from the
    full product, then subtract alpha v and beta v_prev)."""
    n = H.shape[0]
    alpha = np.zeros(niter)
    beta = np.zeros(niter + 1)
    beta[0] = np.linalg.norm(v0)
    q_prev = np.zeros(n)
    r = v0.copy()
    q = np.zeros(n)
    for i in range(1, niter + 1):
        q_prev = q
        q = r / beta[i - 1]
        w = H @ q
        alpha[i - 1] = q @ w
        w = w - alpha[i - 1] * q
        if i > 1:
            w = w - beta[i - 1] * q_prev
        beta[i] = np.linalg.norm(w)
        r = w
    return alpha, beta


def structural_counts(prob):
    """Number of (row, step) matrix elements per unique block and per number
    of distinct modes, counted by brute force over the monomial Kronecker
    operators (all steps, including the zero step of all-even monomials)."""
    ns, nm, nf = prob.nstates, prob.nmodes, list(prob.nfunc)
    blocks = [(i, j) for i in range(ns) for j in range(i + 1)]
    out = np.zeros((prob.ordr, len(blocks)), dtype=np.int64)
    for b in range(len(blocks)):
        for p in range(1, prob.ordr + 1):
            for k, tup in enumerate(mode_tuples(nm, p)):
                c = prob.potterms[b, p - 1, k]
                if not abs(c) >= prob.ztoler:
                    continue
                pows = np.bincount(np.array(tup), minlength=nm)
                # non-zero pattern of q^p restricted to the parity band |i-j| <= p, i-j = p mod 2
                cnt = 1
                for kk in range(nm):
                    if pows[kk] == 0:
                        cnt *= nf[kk]
                        continue
                    ii, jj = np.meshgrid(np.arange(nf[kk]), np.arange(nf[kk]), indexing="ij")
                    d = np.abs(ii - jj)
                    cnt *= int(np.sum((d <= pows[kk]) & ((d - pows[kk]) % 2 == 0)))
                out[int(np.count_nonzero(pows)) - 1, b] += cnt
    return out
