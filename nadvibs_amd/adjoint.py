"""PyTorch autograd over the library's Lanczos run and broadened spectrum.

    env = spectrum_of_parameters(prob, omega, concoef, potterms, niter, grid, sigma)
    loss = torch.linalg.vector_norm(env - target)
    loss.backward()          # d loss / d (omega, concoef, potterms)

spectrum_of_parameters is the path for fits: its backward pass takes
dL/dT from nv_spectrum_adjoint and projects it on the kept Krylov basis
(nv_krylov_gradient), which is stable for any number of iterations.

lanczos_coefficients exposes (alpha, beta) themselves; its backward pass is
the exact reverse sweep of the recurrence (nv_lanczos_adjoint, kept basis or
checkpoints), accurate for short runs only (the sweep amplifies rounding
along converged Ritz directions).

The forward pass runs the FP64 recurrence on the device (H*v in FP64 or
FP32).  Tensors cross the boundary as host float64 arrays: the device work
stays inside the library.  Parameters are the nadvibs.in quantities: aomega
(hartree), concoef[nstblks] and POTterms [nstblks, ordr, noterms(ordr)].
"""
import dataclasses

import numpy as np
import torch

from .engine import Engine
from .inputs import Problem


def _np(t):
    return t.detach().to("cpu", torch.float64).numpy()


class LanczosCoefficients(torch.autograd.Function):
    @staticmethod
    def forward(ctx, omega, concoef, potterms, prob, niter, prec, checkpoint, generic):
        p = dataclasses.replace(prob, aomega=_np(omega).copy(), concoef=_np(concoef).copy(),
                                potterms=_np(potterms).copy(), shiftref=0)
        eng = Engine(p)
        if generic:
            eng.select_hv_kernel(True)
        eng.lanczos_init(niter, reorthog=0, prec=prec, savevecs=checkpoint == 0, checkpoint=checkpoint)
        eng.run(1, niter)
        a, b, _ = eng.coefficients()
        ctx.eng, ctx.prec = eng, prec
        return torch.from_numpy(a), torch.from_numpy(b)

    @staticmethod
    def backward(ctx, ga, gb):
        ga = np.zeros(ctx.eng.niter) if ga is None else _np(ga)
        gb = np.zeros(ctx.eng.niter + 1) if gb is None else _np(gb)
        g = ctx.eng.lanczos_adjoint(ga, gb, ctx.prec)
        ctx.eng.close()
        return (torch.from_numpy(g["omega"]), torch.from_numpy(g["concoef"]), torch.from_numpy(g["potterms"]),
                None, None, None, None, None)


class BroadenedSpectrum(torch.autograd.Function):
    @staticmethod
    def forward(ctx, alpha, beta, grid, sigma, origin, e0, eng):
        a, b, x = _np(alpha), _np(beta), _np(grid)
        ctx.save_for_backward(alpha, beta, grid)
        ctx.args = (sigma, origin, e0, eng)
        return torch.from_numpy(eng.spectrum(a, b, x, sigma, origin, e0))

    @staticmethod
    def backward(ctx, genv):
        alpha, beta, grid = ctx.saved_tensors
        sigma, origin, e0, eng = ctx.args
        ab, bb = eng.spectrum_adjoint(_np(alpha), _np(beta), _np(grid), sigma, _np(genv), origin, e0)
        return torch.from_numpy(ab), torch.from_numpy(bb), None, None, None, None, None


def lanczos_coefficients(prob: Problem, omega, concoef, potterms, niter, prec="f64", checkpoint=0, generic=True):
    """(alpha(1:niter), beta(0:niter)) of the run from the default start vector,
    differentiable in omega, concoef and potterms.  checkpoint = K > 0 keeps
    only every K-th pair of vectors and recomputes the rest in the backward
    pass; generic = True skips the run-time compile of the specialised H*v."""
    return LanczosCoefficients.apply(omega, concoef, potterms, prob, niter, prec, checkpoint, generic)


def broadened_spectrum(alpha, beta, grid, sigma, origin="lowest", e0=0.0, engine=None):
    """Gaussian-broadened sticks (convolve.py:28-48, without the final division
    by the maximum) of the Ritz values and first-component weights of
    T(alpha, beta), on grid (cm^-1); sigma is a standard deviation in cm^-1.
    `engine` supplies the device (any Engine)."""
    if engine is None:
        raise ValueError("broadened_spectrum needs an Engine for its device")
    return BroadenedSpectrum.apply(alpha, beta, grid, sigma, origin, e0, engine)


class SpectrumOfParameters(torch.autograd.Function):
    @staticmethod
    def forward(ctx, omega, concoef, potterms, prob, niter, grid, sigma, origin, e0, prec, reorthog, generic,
                checkpoint):
        p = dataclasses.replace(prob, aomega=_np(omega).copy(), concoef=_np(concoef).copy(),
                                potterms=_np(potterms).copy(), shiftref=0)
        eng = Engine(p)
        if generic:
            eng.select_hv_kernel(True)
        ck = checkpoint if reorthog == 0 else 0
        eng.lanczos_init(niter, reorthog=reorthog, prec=prec, savevecs=ck == 0, checkpoint=ck)
        eng.run(1, niter)
        a, b, _ = eng.coefficients()
        x = _np(grid)
        ctx.eng, ctx.args = eng, (a, b, x, sigma, origin, e0)
        return torch.from_numpy(eng.spectrum(a, b, x, sigma, origin, e0))

    @staticmethod
    def backward(ctx, genv):
        a, b, x, sigma, origin, e0 = ctx.args
        _, _, tb = ctx.eng.spectrum_adjoint(a, b, x, sigma, _np(genv), origin, e0, full=True)
        g = ctx.eng.krylov_gradient(tb)
        ctx.eng.close()
        return (torch.from_numpy(g["omega"]), torch.from_numpy(g["concoef"]), torch.from_numpy(g["potterms"]),
                None, None, None, None, None, None, None, None, None, None)


def spectrum_of_parameters(prob: Problem, omega, concoef, potterms, niter, grid, sigma, origin="lowest", e0=0.0,
                           prec="f64", reorthog=0, generic=True, checkpoint=0):
    """Broadened stick spectrum (convolve.py:28-48 without the division by the
    maximum) of the niter-step run from the default start vector, on grid
    (cm^-1), differentiable in omega, concoef and potterms through the
    Krylov-projected gradient.  reorthog > 0 runs with partial
    re-orthogonalisation; generic = True skips the run-time compile of the
    specialised H*v (each call builds a new operator).  checkpoint = K > 0
    (with reorthog = 0) keeps only (q, r) every K iterations instead of the
    basis; the backward pass then recomputes the basis segment by segment."""
    return SpectrumOfParameters.apply(omega, concoef, potterms, prob, niter, grid, sigma, origin, e0, prec, reorthog,
                                      generic, checkpoint)
