"""Batched Lanczos runs over many Hamiltonians (nv_batch_* in include/nadvibs.h).

    with Batch([prob_a, prob_b, ...]) as bt:
        bt.init(niter=400, prec="f32")
        print(bt.tune())            # rows per workgroup of the batched H*v
        bt.run()
        a, b = bt.coefficients(0)
        ev, inten, bji, sign = bt.ritz(0, a, b)

Every problem advances by one launch sequence per iteration (one workgroup
per tile of rows, over all problems), captured once as a HIP graph.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import NvError, dptr
from .engine import problem_struct


class Batch:
    def __init__(self, probs, device=0):
        self.lib = _lib.load()
        self.probs = list(probs)
        descs, self._keep = [], []
        for p in self.probs:
            d, keep = problem_struct(p)
            descs.append(d)
            self._keep.append(keep)
        arr = (_lib.NvProblem * len(descs))(*descs)
        h = ctypes.c_void_p()
        rc = self.lib.nv_batch_create(arr, len(descs), device, ctypes.byref(h))
        if rc != 0:
            raise NvError(rc, self.lib.nv_batch_last_error(None).decode())
        self.h = h
        self.niter = 0

    def _check(self, rc):
        if rc != 0:
            raise NvError(rc, self.lib.nv_batch_last_error(self.h).decode())

    def init(self, niter, prec="f64"):
        self._check(self.lib.nv_batch_init(self.h, niter, _lib.PREC[prec]))
        self.niter = niter

    def tune(self):
        buf = ctypes.create_string_buffer(2048)
        self._check(self.lib.nv_batch_tune(self.h, buf, 2048))
        return buf.value.decode()

    def run(self, iiter=1, niter=None):
        self._check(self.lib.nv_batch_run(self.h, iiter, niter or self.niter))

    def coefficients(self, p, n=None):
        n = n or self.niter
        a, b = np.zeros(n), np.zeros(n + 1)
        self._check(self.lib.nv_batch_coefficients(self.h, p, dptr(a), dptr(b)))
        return a, b

    def ritz(self, p, alpha, beta):
        """Eigenvalues, intensities, bji and signs of problem p's T."""
        ctx = self.lib.nv_batch_context(self.h, p)
        alpha = np.ascontiguousarray(alpha, dtype=np.float64)
        beta = np.ascontiguousarray(beta, dtype=np.float64)
        n = alpha.size
        ev, it, bj = np.empty(n), np.empty(n), np.empty(n)
        sg = np.empty(n, dtype=np.int32)
        rc = self.lib.nv_ritz(ctx, n, dptr(alpha), dptr(beta), dptr(ev), dptr(it), dptr(bj),
                              sg.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        if rc != 0:
            raise NvError(rc, self.lib.nv_last_error(ctx).decode())
        return ev, it, bj, sg

    def close(self):
        if getattr(self, "h", None):
            self.lib.nv_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
