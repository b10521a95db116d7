"""ctypes wrapper of liboracle.so (TEST INFRASTRUCTURE ONLY).

Mirrors the reference functions restated in nv_oracle.c; see that file for
the citations.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_i64p = ctypes.POINTER(ctypes.c_int64)
_dp = ctypes.POINTER(ctypes.c_double)


class _Problem(ctypes.Structure):
    # layout of nv_problem in include/nadvibs.h
    _fields_ = [
        ("nstates", ctypes.c_int64), ("nmodes", ctypes.c_int64), ("ordr", ctypes.c_int64),
        ("nfunc", _i64p), ("aomega", _dp), ("concoef", _dp), ("potterms", _dp),
        ("ztoler", ctypes.c_double), ("quirks", ctypes.c_int32), ("neworigin", ctypes.c_int32),
        ("statew", _dp), ("bomega", _dp), ("dvec", _dp), ("tmat", _dp),
        ("nirreps", ctypes.c_int64), ("npirrep", _i64p), ("istate", _i64p),
    ]


_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        l = ctypes.CDLL(LIB)
        l.or_create.restype = ctypes.c_void_p
        l.or_create.argtypes = [ctypes.POINTER(_Problem)]
        l.or_destroy.argtypes = [ctypes.c_void_p]
        l.or_dims.restype = ctypes.c_long
        l.or_dims.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_long)]
        l.or_term_counts.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_long)]
        l.or_term_summary.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_long)]
        l.or_hv.argtypes = [ctypes.c_void_p, _dp, _dp]
        l.or_start_vector.argtypes = [ctypes.c_void_p, _dp]
        l.or_lanczos.restype = ctypes.c_int
        l.or_lanczos.argtypes = [ctypes.c_void_p, ctypes.c_long, ctypes.c_int, ctypes.c_long, ctypes.c_uint64,
                                 ctypes.c_double, _dp, _dp, _dp, _dp, _dp, ctypes.POINTER(ctypes.c_long)]
        l.or_hv_rows.argtypes = [ctypes.c_void_p, _dp, ctypes.c_long, ctypes.POINTER(ctypes.c_long), _dp]
        l.or_hv_par.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int]
        l.or_hv_par_range.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_long, ctypes.c_long, ctypes.c_int]
        l.or_lanczos_par.argtypes = [ctypes.c_void_p, ctypes.c_long, ctypes.c_long, ctypes.c_int, _dp, _dp, _dp]
        l.or_set_off32.argtypes = [ctypes.c_int]
        l.or_set_eps_pro.argtypes = [ctypes.c_double]
        l.or_set_pro_thresholds.argtypes = [ctypes.c_double, ctypes.c_double]
        l.or_givens.argtypes = [ctypes.c_long, _dp, _dp, _dp, _dp, _dp, ctypes.POINTER(ctypes.c_int), _dp]
        _lib = l
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp) if a is not None else None


class Oracle:
    """FP64 CPU restatement for one problem (an inputs.Problem-like object)."""

    def __init__(self, prob):
        self.prob = p = prob
        self._keep = [np.ascontiguousarray(p.nfunc, dtype=np.int64),
                      np.ascontiguousarray(p.aomega, dtype=np.float64),
                      np.ascontiguousarray(p.concoef, dtype=np.float64),
                      np.ascontiguousarray(p.potterms, dtype=np.float64),
                      np.ascontiguousarray(p.statew if p.statew is not None else np.ones(p.nstates),
                                           dtype=np.float64)]
        nf, ao, cc, pt, sw = self._keep
        neworigin = int(bool(getattr(p, "shiftref", 0)))
        extra = {}
        if neworigin:
            bo, dv, tm = (np.ascontiguousarray(v, dtype=np.float64) for v in (p.bomega, p.dvec, p.tmat))
            self._keep += [bo, dv, tm]
            extra.update(bomega=_d(bo), dvec=_d(dv), tmat=_d(tm))
        npirr = np.ascontiguousarray(p.npirrep or [p.nmodes], dtype=np.int64)
        ist = np.ascontiguousarray(p.initstate or [0] * p.nmodes, dtype=np.int64)
        self._keep += [npirr, ist]
        desc = _Problem(nstates=p.nstates, nmodes=p.nmodes, ordr=p.ordr, nfunc=nf.ctypes.data_as(_i64p),
                        aomega=_d(ao), concoef=_d(cc), potterms=_d(pt), ztoler=p.ztoler, quirks=p.quirks,
                        neworigin=neworigin, statew=_d(sw), nirreps=len(npirr), npirrep=npirr.ctypes.data_as(_i64p),
                        istate=ist.ctypes.data_as(_i64p), **extra)
        self.h = lib().or_create(ctypes.byref(desc))
        if not self.h:
            raise ValueError("oracle rejected the problem")
        D = ctypes.c_long()
        self.n = lib().or_dims(self.h, ctypes.byref(D))
        self.dimen = D.value

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_destroy(self.h)
            self.h = None

    def hv(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        lib().or_hv(self.h, _d(x), _d(y))
        return y

    def hv_rows(self, x, rows):
        """(H x) at the given rows, all states: out[i, s] = (H x)[s*dimen + rows[i]]."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        assert x.size == self.n and rows.min() >= 0 and rows.max() < self.dimen
        y = np.empty((rows.size, self.prob.nstates))
        lib().or_hv_rows(self.h, _d(x), rows.size, rows.ctypes.data_as(ctypes.POINTER(ctypes.c_long)), _d(y))
        return y

    def hv_par(self, x, threads, rows=None):
        """H x with OpenMP over contiguous row blocks; rows=(r0, r1) computes
        only those rows (the rest of the result is zero)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        if rows is None:
            lib().or_hv_par(self.h, _d(x), _d(y), int(threads))
        else:
            lib().or_hv_par_range(self.h, _d(x), _d(y), int(rows[0]), int(rows[1]), int(threads))
        return y

    def lanczos_par(self, niter, threads, state=None, alpha=None, beta=None, i0=1):
        """Iterations i0..niter of the plain recurrence (reorthog 0) with
        OpenMP threads; pass the returned (state, alpha, beta) back to go on."""
        if state is None:
            state = np.zeros(3 * self.n)
            alpha = np.zeros(niter)
            beta = np.zeros(niter + 1)
        lib().or_lanczos_par(self.h, int(i0), int(niter), int(threads), _d(state), _d(alpha), _d(beta))
        return state, alpha, beta

    def start_vector(self):
        q = np.empty(self.n)
        lib().or_start_vector(self.h, _d(q))
        return q

    def term_counts(self):
        out = np.zeros(self.prob.ordr * self.prob.nstblks, dtype=np.int64)
        lib().or_term_counts(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_long)))
        return out.reshape(self.prob.ordr, self.prob.nstblks)

    def term_summary(self):
        out = np.zeros(self.prob.ordr, dtype=np.int64)
        lib().or_term_summary(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_long)))
        return out

    def lanczos(self, niter, reorthog=0, chkorthog=100, seed=12345, noise=1.0, start=None, keep_basis=False,
                mixed=False, pro_eps=0.0):
        """mixed=True emulates the library's mixed mode: off-diagonal part of
        H*v from a float copy with float accumulation; with the default
        pro_eps (0: 2^-53) exact overlap checks every min(chkorthog, 10)
        iterations and the thresholds MIXED_TAU / MIXED_ETA.  pro_eps > 0 sets
        the roundoff of the omega model (thresholds sqrt(eps), eps^(3/4))."""
        if pro_eps:
            set_eps_pro(pro_eps)
            try:
                return self.lanczos(niter, reorthog, chkorthog, seed, noise, start, keep_basis, mixed)
            finally:
                set_eps_pro(0.0)
        if mixed:
            lib().or_set_off32(1)
            if not lib_eps_set():
                chkorthog = min(chkorthog, 10)
                lib().or_set_pro_thresholds(MIXED_TAU, MIXED_ETA)
            try:
                return self.lanczos(niter, reorthog, chkorthog, seed, noise, start, keep_basis, mixed=False)
            finally:
                lib().or_set_off32(0)
                lib().or_set_pro_thresholds(0.0, 0.0)
        a = np.zeros(niter)
        b = np.zeros(niter + 1)
        om = np.zeros(2 * (niter + 1))
        st = np.zeros(2, dtype=np.int64)
        arch = np.zeros((niter + 1) * self.n) if keep_basis else None
        s = None if start is None else np.ascontiguousarray(start, dtype=np.float64)
        rc = lib().or_lanczos(self.h, niter, reorthog, chkorthog, seed, noise, _d(s), _d(a), _d(b), _d(om),
                              _d(arch), st.ctypes.data_as(ctypes.POINTER(ctypes.c_long)))
        if rc:
            raise MemoryError("oracle Lanczos could not allocate its Krylov basis")
        out = dict(alpha=a, beta=b, omega=om.reshape(2, niter + 1), events=int(st[0]), touched=int(st[1]))
        if keep_basis:
            out["basis"] = arch.reshape(niter + 1, self.n)[1:]
        return out


def set_off32(on):
    """Emulates the library's mixed mode in the oracle: off-diagonal part of
    H*v from a float copy of x with float accumulation."""
    lib().or_set_off32(1 if on else 0)


_EPS_PRO = [0.0]
# NV_PRO_MIXED_TAU / NV_PRO_MIXED_ETA of include/nadvibs.h
MIXED_TAU = 2.0 ** -20
MIXED_ETA = 2.0 ** -36


def set_eps_pro(eps):
    """Unit roundoff of the omega recurrence (0 = the reference's 2^-53)."""
    _EPS_PRO[0] = float(eps)
    lib().or_set_eps_pro(eps)


def lib_eps_set():
    return _EPS_PRO[0] > 0.0


def givens(alpha, beta, vectors=False):
    """GIVENS restated for the tridiagonal T(alpha, beta(1:n-1)); beta holds
    beta(0:n).  Returns (eval, inten, bji, sign[, vecs])."""
    alpha = np.ascontiguousarray(alpha, dtype=np.float64)
    beta = np.ascontiguousarray(beta, dtype=np.float64)
    n = alpha.size
    ev, it, bj = np.empty(n), np.empty(n), np.empty(n)
    sg = np.empty(n, dtype=np.int32)
    vec = np.empty(n * n) if vectors else None
    lib().or_givens(n, _d(alpha), _d(beta), _d(ev), _d(it), _d(bj), sg.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                    _d(vec))
    if vectors:
        return ev, it, bj, sg, vec.reshape(n, n)
    return ev, it, bj, sg


def identify_roots(alpha, beta, basis, nroots, idroots, nstates):
    """Ritz vectors and root analysis restated (identify_roots,
    Source/top.f90:1122-1283): Y_i = sum_k v(k,i) q_k (:1177-1183), the
    idroots largest |Y_i|^2 elements with their signs, ties by increasing
    Fortran index (:1191-1200), state weights in percent (:1203-1210) and the
    overlaps Y_i[:,k] . Y_j[:,l] (:1227-1243).  basis holds q_1..q_n as rows."""
    alpha = np.asarray(alpha, dtype=np.float64)
    n = alpha.size
    ev, inten, bji, sg, vec = givens(alpha, beta, vectors=True)
    Y = vec[:nroots, :n] @ np.asarray(basis)[:n]
    N = Y.shape[1]
    D = N // nstates
    imax = inten[:nroots].max()
    out = dict(eval=ev[:nroots].copy(), inten=inten[:nroots] / imax, sign=sg[:nroots].copy(), vectors=Y)
    ti = np.zeros((nroots, idroots), dtype=np.int64)
    ts = np.zeros((nroots, idroots), dtype=np.int32)
    tp = np.zeros((nroots, idroots), dtype=np.int32)
    tm = np.zeros((nroots, idroots))
    for i in range(nroots):
        m = Y[i] * Y[i]
        order = np.lexsort((np.arange(N), -m))[:idroots]
        ti[i] = order % D + 1
        ts[i] = order // D + 1
        tp[i] = np.where(Y[i, order] < 0, -1, 1)
        tm[i] = m[order]
    out.update(top_index=ti, top_state=ts, top_phase=tp, top_mag2=tm)
    Y3 = Y.reshape(nroots, nstates, D)
    w = np.einsum("ikr,ikr->ik", Y3, Y3)
    out["percent"] = 100.0 * np.abs(w / w.sum(axis=1, keepdims=True))
    out["overlap"] = np.einsum("ikr,jlr->ijkl", Y3, Y3)
    return out


def sticks_envelope(x, h, grid, sigma):
    """Sum of Gaussians h_j exp(-(g - x_j)^2 / (2 sigma^2)) on grid: the
    aperture of convolve.py:28-48 (before its division by the maximum) and of
    the reference's convolve/sub.f90 (sigma = fwhm / (2 sqrt(2 ln 2)))."""
    x, h, grid = (np.asarray(v, dtype=np.float64) for v in (x, h, grid))
    return (h[None, :] * np.exp(-(grid[:, None] - x[None, :]) ** 2 / (2.0 * sigma * sigma))).sum(axis=1)
