"""Seeded synthetic vibronic Hamiltonians for the benchmark workloads.

Couplings are drawn as dimensionless-coordinate coefficients kappa (cm^-1)
and written to nadvibs.in as kappa * prod_k w_k^{p_k/2} (hartree), so that
after the reader's scaling by prod_k w_k^{-p_k/2} (Source/top.f90:212-217)
the coefficient of q-space monomials equals kappa.  Diagonal state blocks
carry a constant (s-1)*gap; inter-state blocks have no constant and no
all-even monomials (so the lower-only diagonal convention of the reference
changes nothing for them) unless `offdiag_even` is set.
"""
import numpy as np

from .inputs import AU2WAVE, Problem, monomials, noterms

# name -> shape and distributions (cm^-1)
WORKLOADS = {
    "cfg1": dict(nstates=2, nfunc=[10, 10, 10], ordr=2, omega=(500, 3000), linear=200.0, quad=50.0,
                 bilinear=50.0, cubic=0.0, quartic=0.0, gap=2000.0, niter=300, reorthog=0, seed=1001),
    "cfg2": dict(nstates=3, nfunc=[12, 10, 10, 8, 8, 6], ordr=2, omega=(300, 3200), linear=300.0, quad=80.0,
                 bilinear=40.0, cubic=0.0, quartic=0.0, gap=2000.0, niter=2000, reorthog=1, chkorthog=100,
                 seed=1002),
    "cfg3": dict(nstates=4, nfunc=[8, 8, 6, 6, 6, 5, 5, 4, 4, 4], ordr=3, omega=(300, 3200), linear=300.0,
                 quad=80.0, bilinear=40.0, cubic=10.0, quartic=0.0, gap=2000.0, niter=3000, reorthog=0,
                 seed=1003),
}


def make_problem(nstates, nfunc, ordr, omega, linear, quad, bilinear, cubic=0.0, quartic=0.0, gap=2000.0,
                 niter=300, reorthog=0, chkorthog=100, seed=0, quirks=3, offdiag_even=0.0, name=""):
    rng = np.random.default_rng(seed)
    nm = len(nfunc)
    w_cm = rng.uniform(omega[0], omega[1], size=nm)
    aomega = w_cm / AU2WAVE
    nb = nstates * (nstates + 1) // 2
    ntmax = noterms(nm, ordr)
    pot = np.zeros((nb, ordr, ntmax))
    concoef = np.zeros(nb)
    blocks = [(i, j) for i in range(nstates) for j in range(i + 1)]
    for b, (i, j) in enumerate(blocks):
        diag = i == j
        if diag:
            concoef[b] = i * gap / AU2WAVE
        for p in range(1, ordr + 1):
            for k, tup in enumerate(monomials(nm, p)):
                pows = np.bincount(np.array(tup), minlength=nm)
                used = pows[pows > 0]
                all_even = bool(np.all(used % 2 == 0))
                if p == 1:
                    sig = linear
                elif p == 2:
                    sig = quad if len(used) == 1 else bilinear
                elif p == 3:
                    sig = cubic
                else:
                    sig = quartic
                kappa = rng.normal(0.0, sig) if sig > 0 else 0.0
                if not diag and all_even:
                    kappa = rng.normal(0.0, offdiag_even) if offdiag_even > 0 else 0.0
                scale = float(np.prod(aomega ** (pows / 2.0)))
                pot[b, p - 1, k] = kappa / AU2WAVE * scale
    return Problem(nstates=nstates, nmodes=nm, ordr=ordr, nfunc=list(nfunc), aomega=aomega, concoef=concoef,
                   potterms=pot, quirks=quirks, niter=niter, reorthog=reorthog, chkorthog=chkorthog,
                   statew=np.ones(nstates), npirrep=[nm], initstate=[0] * nm, name=name)


def workload(name, **over):
    cfg = dict(WORKLOADS[name])
    cfg.update(over)
    return make_problem(name=name, **cfg)


def batch_workloads(count=64, seed=1005):
    """The uneven batch: nstates U{2,3}, nmodes U{4..6}, basis U{4..12}, ordr 2."""
    rng = np.random.default_rng(seed)
    out = []
    for b in range(count):
        ns = int(rng.integers(2, 4))
        nm = int(rng.integers(4, 7))
        nf = [int(v) for v in rng.integers(4, 13, size=nm)]
        out.append(make_problem(ns, nf, 2, (300, 3200), 300.0, 80.0, 40.0, niter=400,
                                seed=int(rng.integers(1 << 30)), name=f"batch{b}"))
    return out
