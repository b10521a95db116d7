"""Ritz vectors and root analysis (identify_roots, Source/top.f90:1122-1283).

The oracle's restatement (oracle.oracle.identify_roots) is pinned against
the dense eigenvectors of H: converged Ritz vectors equal +-psi_i, so their
state weights, largest components and mutual overlaps follow from
numpy.linalg.eigh.  The HIP path is then checked against the oracle on the
same Krylov basis (the library's own, read back through the C ABI): Ritz
vectors and overlaps to 1e-12 relative, selected indices exactly.
"""
import os

import numpy as np
import pytest

from nadvibs_amd.engine import Engine, barray, rootinfo_text, sodata_text
from nadvibs_amd.synth import make_problem, workload
from oracle.dense import dense_hamiltonian
from oracle.oracle import Oracle, givens, identify_roots


def _small(ns=2, nf=(6, 5, 4), seed=5):
    return make_problem(ns, list(nf), 2, (500, 3000), 250.0, 60.0, 40.0, seed=seed)


def test_oracle_roots_match_dense_eigenvectors():
    p = _small()
    H = dense_hamiltonian(p)
    w, U = np.linalg.eigh(H)
    r = Oracle(p).lanczos(120, reorthog=1, keep_basis=True)
    nroots, idroots = 4, 3
    res = identify_roots(r["alpha"], r["beta"], r["basis"], nroots, idroots, p.nstates)
    np.testing.assert_allclose(res["eval"], w[:nroots], rtol=1e-12, atol=1e-14)
    D = p.dimen
    for i in range(nroots):
        Y = res["vectors"][i]
        s = np.sign(Y @ U[:, i])
        np.testing.assert_allclose(s * Y, U[:, i], atol=1e-8)
        # state weights
        wts = (U[:, i].reshape(p.nstates, D) ** 2).sum(axis=1)
        np.testing.assert_allclose(res["percent"][i], 100 * wts / wts.sum(), atol=1e-6)
        # largest components of the exact eigenvector, 1-based (row, state)
        m = U[:, i] ** 2
        top = np.argsort(-m, kind="stable")[:idroots]
        assert np.all(np.diff(m[np.argsort(-m)][: idroots + 1]) < -1e-9), "near-tie, pick another seed"
        np.testing.assert_array_equal(res["top_index"][i], top % D + 1)
        np.testing.assert_array_equal(res["top_state"][i], top // D + 1)
        np.testing.assert_array_equal(res["top_phase"][i], np.where(s * U[top, i] < 0, -1, 1))
        np.testing.assert_allclose(res["top_mag2"][i], m[top], atol=1e-8)
    # overlaps: blocks of the eigenvectors, with the Ritz vectors' signs
    U3 = U[:, :nroots].T.reshape(nroots, p.nstates, D)
    sg = np.array([np.sign(res["vectors"][i] @ U[:, i]) for i in range(nroots)])
    ref = np.einsum("ikr,jlr->ijkl", U3, U3) * sg[:, None, None, None] * sg[None, :, None, None]
    np.testing.assert_allclose(res["overlap"], ref, atol=1e-8)
    # diagonal blocks sum to the squared norm (1 for an eigenvector)
    for i in range(nroots):
        assert abs(np.trace(res["overlap"][i, i]) - 1.0) < 1e-8


def test_barray_decodes_fortran_index():
    """get_barray (Source/secondary.f90:812-824): last mode fastest."""
    p = _small()
    assert barray(p, 1) == [0, 0, 0]
    assert barray(p, p.dimen) == [f - 1 for f in p.nfunc]
    assert barray(p, 1 + 2 * 5 * 4 + 3 * 4 + 1) == [2, 3, 1]


def test_rootinfo_and_sodata_formats():
    """Layout of rootinfo.dat and sodata.in (Source/top.f90:1258-1280)."""
    p = _small(ns=3)
    res = dict(eval=np.array([0.01, 0.012]), inten=np.array([1.0, 0.25]), sign=np.array([1, -1], dtype=np.int32),
               top_index=np.array([[1, 7], [2, 3]]), top_state=np.array([[1, 2], [3, 1]], dtype=np.int32),
               top_phase=np.array([[1, -1], [-1, 1]], dtype=np.int32),
               top_mag2=np.array([[0.5, 0.25], [0.4, 0.1]]), percent=np.array([[50.0, 30.0, 20.0], [1, 2, 97.0]]),
               overlap=np.arange(36, dtype=float).reshape(2, 2, 3, 3) * 1e-3)
    txt = rootinfo_text(p, res).splitlines()
    assert txt[0] == " Summary for computing    2 eigenvectors corresponding to the lowest eigenvalues"
    assert txt[3] == " ************ ROOT    1 *************"
    assert txt[4] == " Energy:   2194.75 cm-1,   0.27211 eV   Intensity: 1.0000"
    assert txt[6] == " State vector:  1 | Index:            1 Phase:  1 | Magnitude^2: 0.500000"
    assert txt[7] == " v(i):   0  0  0"
    assert txt[9] == " v(i):   0  1  2"
    assert txt[11] == " Percent State (1,2,3): ( 50.00 %,  30.00 %,  20.00 %)"
    assert txt[-1] == "------------ FINISHED EXECUTION ------------"
    so = sodata_text(res, 2).splitlines()
    assert so[0] == "    2"
    assert so[3] == "    2   3.26536614E-01  -5.00000000E-01"
    # root 2 with root 1: [S1(2)S2(1)] = overlap[1, 0, 0, 1], then (0,2), (1,0) | (1,2), (2,0), (2,1)
    assert so[6] == "    2     1    1.900000E-02   2.000000E-02   2.100000E-02"
    assert so[7] == "    2     1    2.300000E-02   2.400000E-02   2.500000E-02"


def _gpu_roots(p, niter, nroots, idroots, reorthog=0):
    with Engine(p) as eng:
        eng.lanczos_init(niter, reorthog=reorthog, savevecs=True)
        eng.run(1, niter)
        a, b, _ = eng.coefficients()
        res = eng.identify_roots(a, b, nroots, idroots, vectors=True)
        basis = np.array([eng.lanczos_vector(j) for j in range(1, niter + 1)])
    return a, b, basis, res


@pytest.mark.gpu
@pytest.mark.parametrize("shape,niter,nroots,idroots,reorthog", [
    ((2, (6, 5, 4)), 60, 5, 4, 0),
    ((3, (10, 10, 10, 8)), 40, 12, 6, 0),
    ((3, (4, 3, 5)), 50, 7, 3, 1),
], ids=["small", "multi-tile", "pro-archive"])
def test_identify_roots_gpu_matches_oracle(shape, niter, nroots, idroots, reorthog):
    ns, nf = shape
    p = _small(ns, nf, seed=9)
    a, b, basis, res = _gpu_roots(p, niter, nroots, idroots, reorthog)
    ref = identify_roots(a, b, basis, nroots, idroots, ns)
    scale = np.max(np.abs(ref["vectors"]))
    assert np.max(np.abs(res["vectors"] - ref["vectors"])) <= 1e-12 * scale
    np.testing.assert_allclose(res["eval"], ref["eval"], rtol=1e-13, atol=1e-16)
    np.testing.assert_allclose(res["inten"], ref["inten"], rtol=1e-10, atol=1e-14)
    np.testing.assert_array_equal(res["sign"], ref["sign"])
    np.testing.assert_allclose(res["percent"], ref["percent"], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(res["overlap"], ref["overlap"], rtol=1e-10, atol=1e-12)
    # the selection reads the library's own vectors: indices must agree exactly
    # except where two squared magnitudes tie to rounding
    for i in range(nroots):
        m = res["vectors"][i] ** 2
        order = np.lexsort((np.arange(m.size), -m))[:idroots]
        np.testing.assert_array_equal(res["top_index"][i], order % p.dimen + 1)
        np.testing.assert_array_equal(res["top_state"][i], order // p.dimen + 1)
        np.testing.assert_array_equal(res["top_phase"][i], np.where(res["vectors"][i, order] < 0, -1, 1))
        np.testing.assert_array_equal(res["top_mag2"][i], m[order])
        np.testing.assert_array_equal(res["top_index"][i], ref["top_index"][i])


@pytest.mark.gpu
def test_identify_roots_needs_the_archive():
    p = _small()
    with Engine(p) as eng:
        eng.lanczos_init(20)
        eng.run(1, 20)
        a, b, _ = eng.coefficients()
        with pytest.raises(Exception, match="savevecs"):
            eng.identify_roots(a, b, 2, 1)


@pytest.mark.gpu
def test_driver_writes_rootinfo_and_sodata(tmp_path):
    """The Fortran driver with idroots/soroots writes rootinfo.dat and sodata.in
    whose numbers match the Python binding on the same run."""
    from test_driver import _build, run_driver
    from nadvibs_amd.inputs import write_problem
    _build()
    p = workload("cfg1")
    p.niter = 120
    p.idroots, p.soroots = 3, 3
    write_problem(tmp_path, p)
    run_driver(tmp_path)
    with Engine(p) as eng:
        eng.lanczos_init(p.niter, savevecs=True)
        eng.run(1, p.niter)
        a, b, _ = eng.coefficients()
        res = eng.identify_roots(a, b, 3, 3)
    got = open(tmp_path / "rootinfo.dat").read().splitlines()
    want = rootinfo_text(p, res).splitlines()
    assert len(got) == len(want)
    for g, w in zip(got, want):
        assert g.split() == w.split() or _close(g, w), (g, w)
    got = open(tmp_path / "sodata.in").read().splitlines()
    want = sodata_text(res, 3).splitlines()
    assert len(got) == len(want)
    for g, w in zip(got, want):
        assert g.split() == w.split() or _close(g, w), (g, w)


def _close(a, b):
    ta, tb = a.split(), b.split()
    if len(ta) != len(tb):
        return False
    for x, y in zip(ta, tb):
        if x == y:
            continue
        try:
            if abs(float(x.rstrip(",%)")) - float(y.rstrip(",%)"))) > 1e-6 * max(1.0, abs(float(y.rstrip(",%)")))):
                return False
        except ValueError:
            return False
    return True
