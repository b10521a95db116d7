"""PyTorch autograd over the library's Lanczos run and broadened spectrum.

    alpha, beta = lanczos_coefficients(prob, omega, concoef, potterms, niter)
    env = broadened_spectrum(alpha, beta, grid, sigma)
    loss = torch.linalg.vector_norm(env - target)
    loss.backward()          # d loss / d (omega, concoef, potterms)

The forward pass runs the FP64 recurrence on the device (H*v in FP64 or
FP32); the backward pass is the library's hand-written reverse sweep
(nv_lanczos_adjoint), which reads the kept Krylov basis or recomputes it
from checkpoints, and nv_spectrum_adjoint for the spectrum.  Tensors cross
the boundary as host float64 arrays: the device work stays inside the
library.  Parameters are the nadvibs.in quantities: aomega (hartree),
concoef[nstblks] and POTterms [nstblks, ordr, noterms(ordr)].
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
