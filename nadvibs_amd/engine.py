"""Python host side of the engine: mirrors the reference's Lanczos driver
(Source/main.f90:57-78) on top of the C ABI.

    eng = Engine(problem)
    eng.lanczos_init(niter=300)
    for i in range(1, 301):
        eng.update_lanczos(i); eng.Hv(i); eng.lanczos_step(i)
        if orthog and i > 1: eng.set_omega(i); eng.partial_orthog(i)
    alpha, beta, _ = eng.coefficients()
    table = compute_eigenvalues(eng, alpha, beta, bconv)
"""
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import NvError, dptr, iptr
from .inputs import AU2EV, AU2WAVE, Problem


def problem_struct(p: Problem):
    """The nv_problem of p and the arrays it points into (keep them alive)."""
    nfunc = np.ascontiguousarray(p.nfunc, dtype=np.int64)
    aomega = np.ascontiguousarray(p.aomega, dtype=np.float64)
    concoef = np.ascontiguousarray(p.concoef, dtype=np.float64)
    pot = np.ascontiguousarray(p.potterms, dtype=np.float64)
    statew = np.ascontiguousarray(p.statew if p.statew is not None else np.ones(p.nstates), dtype=np.float64)
    npirr = np.ascontiguousarray(p.npirrep or [p.nmodes], dtype=np.int64)
    istate = np.ascontiguousarray(p.initstate or [0] * p.nmodes, dtype=np.int64)
    neworigin = bool(p.shiftref)
    bo = dv = tm = None
    if neworigin:
        bo = np.ascontiguousarray(p.bomega, dtype=np.float64)
        dv = np.ascontiguousarray(p.dvec, dtype=np.float64)
        tm = np.ascontiguousarray(p.tmat, dtype=np.float64)
    keep = [nfunc, aomega, concoef, pot, statew, npirr, istate, bo, dv, tm]
    desc = _lib.NvProblem(
        nstates=p.nstates, nmodes=p.nmodes, ordr=p.ordr, nfunc=iptr(nfunc), aomega=dptr(aomega),
        concoef=dptr(concoef), potterms=dptr(pot), ztoler=p.ztoler, quirks=p.quirks, neworigin=int(neworigin),
        statew=dptr(statew), bomega=dptr(bo), dvec=dptr(dv), tmat=dptr(tm), nirreps=len(npirr),
        npirrep=iptr(npirr), istate=iptr(istate))
    return desc, keep


class Engine:
    def __init__(self, prob: Problem, device=0):
        self.lib = _lib.load()
        self.prob = prob
        desc, self._keep = problem_struct(prob)
        h = ctypes.c_void_p()
        rc = self.lib.nv_create(ctypes.byref(desc), device, ctypes.byref(h))
        if rc != 0:
            raise NvError(rc, self.lib.nv_last_error(None).decode())
        self.h = h
        d, n = ctypes.c_int64(), ctypes.c_int64()
        self.lib.nv_dims(h, ctypes.byref(d), ctypes.byref(n))
        self.dimen, self.n = d.value, n.value
        self.niter = 0

    # -- plumbing ---------------------------------------------------------
    def _check(self, rc):
        if rc != 0:
            raise NvError(rc, self.lib.nv_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.nv_destroy(self.h)
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

    # -- operator -----------------------------------------------------------
    def term_summary(self):
        out = np.zeros(self.prob.ordr, dtype=np.int64)
        self._check(self.lib.nv_term_summary(self.h, iptr(out)))
        return out

    def term_counts(self):
        """ntcom(blk, i) of the reference's check, shape [ordr, nstblks]."""
        out = np.zeros(self.prob.ordr * self.prob.nstblks, dtype=np.int64)
        self._check(self.lib.nv_term_counts(self.h, iptr(out)))
        return out.reshape(self.prob.ordr, self.prob.nstblks)

    def hv(self, x, prec="f64"):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        if x.size != self.n:
            raise ValueError("x has the wrong length")
        y = np.empty_like(x)
        self._check(self.lib.nv_hv(self.h, _lib.PREC[prec], dptr(x), dptr(y)))
        return y

    def hv_device(self, x_ptr, y_ptr, prec="f64", stream=None):
        self._check(self.lib.nv_hv_device(self.h, _lib.PREC[prec], ctypes.c_void_p(x_ptr), ctypes.c_void_p(y_ptr),
                                          ctypes.c_void_p(stream) if stream else None))

    def start_vector(self):
        q = np.empty(self.n)
        self._check(self.lib.nv_start_vector(self.h, dptr(q)))
        return q

    # -- Lanczos ---------------------------------------------------------------
    def lanczos_init(self, niter, reorthog=0, prec="f64", chkorthog=100, seed=12345, noise=1.0, pro_eps=0.0,
                     savevecs=False, checkpoint=0, pro_tau=0.0, pro_eta=0.0):
        o = _lib.NvLanczosOpts(niter=niter, reorthog=reorthog, prec=_lib.PREC[prec], chkorthog=chkorthog,
                               seed=seed, noise_scale=noise, pro_eps=pro_eps, savevecs=1 if savevecs else 0,
                               checkpoint=checkpoint, pro_tau=pro_tau, pro_eta=pro_eta)
        self._check(self.lib.nv_lanczos_init(self.h, ctypes.byref(o)))
        self.niter = niter
        self.reorthog = reorthog

    def update_lanczos(self, i):
        self._check(self.lib.nv_update_lanczos(self.h, i))

    def Hv(self, i):
        self._check(self.lib.nv_hv_iter(self.h, i))

    def lanczos_step(self, i):
        self._check(self.lib.nv_lanczos_step(self.h, i))

    def set_omega(self, i):
        self._check(self.lib.nv_set_omega(self.h, i))

    def partial_orthog(self, i):
        self._check(self.lib.nv_partial_orthog(self.h, i))

    def run(self, iiter=1, niter=None):
        self._check(self.lib.nv_lanczos_run(self.h, iiter, niter or self.niter))

    def coefficients(self, n=None, omega=False):
        n = n or self.niter
        a = np.zeros(n)
        b = np.zeros(n + 1)
        o = np.zeros(2 * (n + 1)) if omega else None
        self._check(self.lib.nv_get_coefficients(self.h, dptr(a), dptr(b), dptr(o)))
        return a, b, (o.reshape(n + 1, 2).T.copy() if omega else None)

    def pro_stats(self):
        e, t = ctypes.c_int64(), ctypes.c_int64()
        self._check(self.lib.nv_pro_stats(self.h, ctypes.byref(e), ctypes.byref(t)))
        return e.value, t.value

    def lanczos_vector(self, j):
        q = np.empty(self.n)
        self._check(self.lib.nv_get_lanczos_vector(self.h, j, dptr(q)))
        return q

    def ritz(self, alpha, beta):
        """Eigenvalues, intensities, bji and signs of T (alpha, beta(0:n))."""
        alpha = np.ascontiguousarray(alpha, dtype=np.float64)
        beta = np.ascontiguousarray(beta, dtype=np.float64)
        n = alpha.size
        if beta.size != n + 1:
            raise ValueError("beta must hold beta(0:n)")
        ev, it, bj = np.empty(n), np.empty(n), np.empty(n)
        sg = np.empty(n, dtype=np.int32)
        self._check(self.lib.nv_ritz(self.h, n, dptr(alpha), dptr(beta), dptr(ev), dptr(it), dptr(bj),
                                     sg.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return ev, it, bj, sg

    def identify_roots(self, alpha, beta, nroots, idroots=0, vectors=False):
        """Ritz vectors of the lowest nroots roots and their analysis
        (identify_roots, Source/top.f90:1122-1283).  Needs the Krylov basis
        (lanczos_init with savevecs or reorthog).  Returns a dict with eval,
        inten (relative to the strongest of the nroots), sign, top_index
        (1-based vibrational index), top_state (1-based), top_phase, top_mag2
        ([nroots, idroots]), percent [nroots, nstates], overlap
        [nroots, nroots, nstates, nstates] and, if asked, vectors [nroots, N]."""
        alpha = np.ascontiguousarray(alpha, dtype=np.float64)
        beta = np.ascontiguousarray(beta, dtype=np.float64)
        n = alpha.size
        if beta.size != n + 1:
            raise ValueError("beta must hold beta(0:n)")
        ns = self.prob.nstates
        k = max(idroots, 1)
        out = dict(eval=np.empty(nroots), inten=np.empty(nroots), sign=np.empty(nroots, dtype=np.int32),
                   top_index=np.zeros((nroots, k), dtype=np.int64), top_state=np.zeros((nroots, k), dtype=np.int32),
                   top_phase=np.zeros((nroots, k), dtype=np.int32), top_mag2=np.zeros((nroots, k)),
                   percent=np.empty((nroots, ns)), overlap=np.empty((nroots, nroots, ns, ns)))
        vec = np.empty((nroots, self.n)) if vectors else None
        i32 = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        self._check(self.lib.nv_identify_roots(
            self.h, n, dptr(alpha), dptr(beta), nroots, idroots, dptr(out["eval"]), dptr(out["inten"]),
            i32(out["sign"]), out["top_index"].ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), i32(out["top_state"]),
            i32(out["top_phase"]), dptr(out["top_mag2"]), dptr(out["percent"]), dptr(out["overlap"]), dptr(vec)))
        for key in ("top_index", "top_state", "top_phase", "top_mag2"):
            out[key] = out[key][:, :idroots]
        if vectors:
            out["vectors"] = vec
        return out

    # -- derivatives -----------------------------------------------------------
    def select_hv_kernel(self, generic=True):
        """generic=True skips the run-time compile of the specialised H*v;
        generic="steps" also keeps the generic kernel in its step-table form."""
        code = 2 if generic == "steps" else (1 if generic else 0)
        self._check(self.lib.nv_select_hv_kernel(self.h, code))

    def lanczos_adjoint(self, abar, bbar, prec="f64"):
        """dL/d(aomega, concoef, potterms) of the last run, given
        abar = dL/dalpha(1:n) and bbar = dL/dbeta(0:n)."""
        abar = np.ascontiguousarray(abar, dtype=np.float64)
        bbar = np.ascontiguousarray(bbar, dtype=np.float64)
        n = abar.size
        if bbar.size != n + 1:
            raise ValueError("bbar must hold beta(0:n)")
        p = self.prob
        go, gc = np.zeros(p.nmodes), np.zeros(p.nstblks)
        gp = np.zeros(np.shape(p.potterms))
        self._check(self.lib.nv_lanczos_adjoint(self.h, n, dptr(abar), dptr(bbar), _lib.PREC[prec], dptr(go),
                                                dptr(gc), dptr(gp)))
        return dict(omega=go, concoef=gc, potterms=gp)

    def spectrum(self, alpha, beta, grid, sigma, origin="lowest", e0=0.0):
        """Broadened sticks of T(alpha, beta) on grid (cm^-1), origin theta_1
        ("lowest") or e0 hartree ("fixed")."""
        alpha, beta, grid = (np.ascontiguousarray(v, dtype=np.float64) for v in (alpha, beta, grid))
        env = np.empty_like(grid)
        self._check(self.lib.nv_spectrum(self.h, alpha.size, dptr(alpha), dptr(beta), grid.size, dptr(grid),
                                         float(sigma), 0 if origin == "lowest" else 1, float(e0), dptr(env)))
        return env

    def spectrum_adjoint(self, alpha, beta, grid, sigma, gbar, origin="lowest", e0=0.0, full=False):
        """(abar, bbar) of L with dL/denv = gbar; with full=True also dL/dT
        as a symmetric n x n matrix."""
        alpha, beta, grid, gbar = (np.ascontiguousarray(v, dtype=np.float64) for v in (alpha, beta, grid, gbar))
        n = alpha.size
        ab, bb = np.empty(n), np.empty(n + 1)
        tb = np.empty((n, n)) if full else None
        self._check(self.lib.nv_spectrum_adjoint(self.h, n, dptr(alpha), dptr(beta), grid.size, dptr(grid),
                                                 float(sigma), 0 if origin == "lowest" else 1, float(e0), dptr(gbar),
                                                 dptr(ab), dptr(bb), dptr(tb)))
        return (ab, bb, tb) if full else (ab, bb)

    def krylov_gradient(self, tbar, strict=False):
        """dL/d(aomega, concoef, potterms) of a spectral loss with dL/dT = tbar,
        projected on the kept Krylov basis (nv_krylov_gradient).  A basis that
        has lost orthogonality (status NV_WARN_NONORTHOGONAL) gives a
        RuntimeWarning naming the overlap, or NvError when strict; the message
        is kept in self.orthogonality_warning (None when the basis passed).
        Such a projection can still match the derivative of the computed loss
        (short runs whose spectrum has converged), but nothing guarantees it."""
        tbar = np.ascontiguousarray(tbar, dtype=np.float64)
        n = tbar.shape[0]
        p = self.prob
        go, gc = np.zeros(p.nmodes), np.zeros(p.nstblks)
        gp = np.zeros(np.shape(p.potterms))
        rc = self.lib.nv_krylov_gradient(self.h, n, dptr(tbar), dptr(go), dptr(gc), dptr(gp))
        self.orthogonality_warning = None
        if rc == _lib.NV_WARN_NONORTHOGONAL and not strict:
            import warnings
            self.orthogonality_warning = self.lib.nv_last_error(self.h).decode()
            warnings.warn(self.orthogonality_warning, RuntimeWarning)
        else:
            self._check(rc)
        return dict(omega=go, concoef=gc, potterms=gp)

    def hv_kernel(self, prec="f64"):
        """Which H*v kernel serves `prec` (compiles it on first use)."""
        buf = ctypes.create_string_buffer(256)
        self._check(self.lib.nv_hv_kernel_info(self.h, _lib.PREC[prec], buf, 256))
        return buf.value.decode()

    def convolve(self, x, y, grid, sigma):
        """Normalised Gaussian envelope of sticks (x, y) on `grid` (convolve.py:28-48)."""
        x, y, grid = (np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, grid))
        env = np.empty_like(grid)
        self._check(self.lib.nv_convolve(self.h, x.size, dptr(x), dptr(y), grid.size, dptr(grid), float(sigma),
                                         dptr(env)))
        return env

    def profile(self, on=True):
        self._check(self.lib.nv_profile_enable(self.h, 1 if on else 0))

    def hv_flops(self):
        """Floating-point operations of one H*v (2 x nonzero multiply-adds)."""
        f = ctypes.c_int64()
        self._check(self.lib.nv_hv_flops(self.h, ctypes.byref(f)))
        return f.value

    def profile_hv(self, launches=200):
        """Average microseconds of back-to-back iteration H*v launches (after update_lanczos(1))."""
        us = ctypes.c_double()
        self._check(self.lib.nv_profile_hv(self.h, launches, ctypes.byref(us)))
        return us.value

    def profile_read(self):
        """(milliseconds, launches) of the timed H*v launches since the last read."""
        ms, n = ctypes.c_double(), ctypes.c_int64()
        self._check(self.lib.nv_profile_read(self.h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def save_vectors(self, path):
        self._check(self.lib.nv_save_vectors(self.h, os.fsencode(path)))

    def load_vectors(self, path, it, alpha, beta, omega=None):
        """omega: [2, it+1] rows omega(1, 0:it), omega(2, 0:it) of restart.log, or None."""
        alpha = np.ascontiguousarray(alpha, dtype=np.float64)
        beta = np.ascontiguousarray(beta, dtype=np.float64)
        om = None
        if omega is not None:
            om = np.ascontiguousarray(np.asarray(omega, dtype=np.float64)[:, :it + 1].T)  # [j][row]
        self._check(self.lib.nv_load_vectors(self.h, os.fsencode(path), it, dptr(alpha), dptr(beta),
                                             None if om is None else dptr(om)))

    def save_archive(self, path, pro_path=None):
        """The kept basis q_1..q_n (and the re-orthogonalisation state)."""
        self._check(self.lib.nv_save_archive(self.h, os.fsencode(path),
                                             None if pro_path is None else os.fsencode(pro_path)))

    def load_archive(self, path, it, pro_path=None):
        self._check(self.lib.nv_load_archive(self.h, os.fsencode(path),
                                             None if pro_path is None else os.fsencode(pro_path), it))


def select_roots(evals, inten, bji, bconv=0):
    """Roots kept by compute_eigenvalues (Source/top.f90:1066-1086): the
    lowest root and every later root with bji < 10^-bconv (the merging
    threshold there is 0, so nothing merges).  Returns 0-based indices."""
    bjiconv = 10.0 ** (-bconv)
    keep = [0]
    for i in range(1, len(evals)):
        if bji[i] < bjiconv:
            keep.append(i)
    return np.array(keep, dtype=np.int64)


def _es13_4(v):
    s = "%.4E" % v
    mant, ex = s.split("E")
    e = int(ex)
    s = f"{mant}E{'-' if e < 0 else '+'}{abs(e):02d}" if abs(e) < 100 else f"{mant}{'-' if e < 0 else '+'}{abs(e):03d}"
    return s.rjust(13)


def eigen_table(prob, evals, inten, bji, bconv=0):
    """The output.dat section written by compute_eigenvalues
    (Source/top.f90:1057-1112), as text."""
    keep = select_roots(evals, inten, bji, bconv)
    zpve = sum(0.5 * w for w in prob.aomega)
    imax = max(inten[k] for k in keep)
    lines = ["", "COMPUTED EIGENVALUES ", "------------------------------", "",
             "  ZPVE from nadvibs.in (reference state): %12.4f cm-1" % (zpve * AU2WAVE),
             "  FIRST eigenvalue:                       %12.4f cm-1" % (evals[keep[0]] * AU2WAVE),
             "  Intensity Factor (MAX Intensity):       %12.8f" % imax, "",
             "  Root  Index  Convergence     E(cm-1)        E(eV)     Intensity"]
    for r, k in enumerate(keep):
        lines.append("%6d%7d%s%15.5f%12.5f%12.7f" % (r + 1, k + 1, _es13_4(bji[k]),
                                                     (evals[k] - evals[keep[0]]) * AU2WAVE,
                                                     (evals[k] - evals[0]) * AU2EV, inten[k] / imax))
    lines += ["", "   All requested eigenvalues computed."]
    return "\n".join(lines) + "\n"


def _es(v, w, d):
    """Fortran ESw.d."""
    s = "%.*E" % (d, v)
    mant, ex = s.split("E")
    e = int(ex)
    s = f"{mant}E{'-' if e < 0 else '+'}{abs(e):02d}" if abs(e) < 100 else f"{mant}{'-' if e < 0 else '+'}{abs(e):03d}"
    return s.rjust(w)


def barray(prob, index):
    """0-based quantum numbers of a 1-based vibrational index (get_barray,
    Source/secondary.f90:812-824; the last mode varies fastest)."""
    lab = max(int(index) - 1, 0)
    out = [0] * prob.nmodes
    for m in range(prob.nmodes - 1, -1, -1):
        out[m] = lab % prob.nfunc[m]
        lab //= prob.nfunc[m]
    return out


def rootinfo_text(prob, res):
    """rootinfo.dat as identify_roots writes it (Source/top.f90:1158-1225,
    formats :1266-1278), from the dict of Engine.identify_roots."""
    nroots, ns = res["percent"].shape
    idroots = res["top_index"].shape[1]
    out = [" Summary for computing %4d eigenvectors corresponding to the lowest eigenvalues" % nroots]
    for i in range(nroots):
        out += [" Eigenvector %4d has been computed..." % (i + 1), "", " ************ ROOT %4d *************" % (i + 1),
                " Energy: %9.2f cm-1, %9.5f eV   Intensity: %6.4f" % (res["eval"][i] * AU2WAVE,
                                                                     res["eval"][i] * AU2EV, res["inten"][i]), ""]
        for j in range(idroots):
            out.append(" State vector: %2d | Index: %12d Phase: %2d | Magnitude^2: %8.6f" % (
                res["top_state"][i, j], res["top_index"][i, j], res["top_phase"][i, j], res["top_mag2"][i, j]))
            out.append(" v(i): " + "".join("%3d" % b for b in barray(prob, res["top_index"][i, j])))
        out.append(" ")
        if 2 <= ns <= 5:
            lab = ",".join(str(k + 1) for k in range(ns))
            out.append(" Percent State (%s): (" % lab + ", ".join("%6.2f %%" % v for v in res["percent"][i]) + ")")
        out.append(" --------------------------------------------------------------")
    out += [" ", "------------ FINISHED EXECUTION ------------"]
    return "\n".join(out) + "\n"


def sodata_text(res, soroots):
    """sodata.in as identify_roots writes it (Source/top.f90:1169-1179,
    :1227-1243, formats :1260-1265, :1279-1280)."""
    ns = res["percent"].shape[1]
    out = ["%5d" % soroots, " Root   Energy (eV)      Sqrt(Int.)*sign"]
    for i in range(soroots):
        out.append("%5d  %s  %s" % (i + 1, _es(res["eval"][i] * AU2EV, 15, 8),
                                    _es(np.sqrt(res["inten"][i]) * res["sign"][i], 15, 8)))
    if ns == 2:
        out.append(" R(i)  R(j)   [S1(i)S2(j)]   [S2(i)S1(j)]")
    if ns == 3:
        out += [" R(i)  R(j)   [S1(i)S2(j)]   [S1(i)S3(j)]  [S2(i)S1(j)]",
                " R(i)  R(j)   [S2(i)S3(j)]   [S3(i)S1(j)]  [S3(i)S2(j)]"]
    for i in range(soroots):
        for j in range(i):
            dp = [res["overlap"][i, j, k, l] for k in range(ns) for l in range(ns) if k != l]
            if ns == 2:
                out.append("%5d %5d  %s %s" % (i + 1, j + 1, _es(dp[0], 14, 6), _es(dp[1], 14, 6)))
            if ns == 3:
                for h in (0, 3):
                    out.append("%5d %5d  %s %s %s" % (i + 1, j + 1, _es(dp[h], 14, 6), _es(dp[h + 1], 14, 6),
                                                      _es(dp[h + 2], 14, 6)))
    return "\n".join(out) + "\n"


def write_restart_log(path, alpha, beta, omega=None):
    """restart.log as make_restart writes it (Source/top.f90:1020-1039)."""
    n = len(alpha)
    fmt = lambda a: " " + " ".join("%.17E" % v for v in a)
    out = [" TOTAL NUMBER OF ITERATIONS IN FILE", " %d" % n, " ITERATION NUMBER TO RESTART FROM", " %d" % n,
           " ALPHA COEFFICIENTS -- diagonal T matrix elements", fmt(alpha),
           " BETA COEFFICIENTS -- off-diagonal T matrix elements", fmt(beta)]
    if omega is not None:
        out += [" OMEGA(1) -- ORTHOGONALITY INFORMATION", fmt(omega[0]),
                " OMEGA(2) -- ORTHOGONALITY INFORMATION", fmt(omega[1])]
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def read_restart_log(path, orthog=False):
    """Reads restart.log (Source/secondary.f90:872-892): returns (total,
    restart_iter, alpha(1:n), beta(0:n), omega or None)."""
    from .inputs import _ListReader
    with open(path) as f:
        rd = _ListReader(f.read().splitlines())
    rd.skip(); tot = int(rd.read(1)[0])
    rd.skip(); n = int(rd.read(1)[0])
    rd.skip(); a = rd.read(n)
    rd.skip(); b = rd.read(n + 1)
    om = None
    if orthog:
        rd.skip(); o1 = rd.read(n + 1)
        rd.skip(); o2 = rd.read(n + 1)
        om = np.vstack([o1, o2])
    return tot, n, a, b, om
