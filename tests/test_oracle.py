"""Pins the oracle (oracle/nv_oracle.c) before it is trusted as the checker.

The reference ships no tests or fixtures, and Global Arrays is absent so it
cannot be built here; each oracle function is pinned by mathematics instead:
  * H*v      -- dense Kronecker brute force from ladder-operator algebra
                (oracle/dense.py), both operator conventions, ordr 1..4;
                zero-coupling and displaced-oscillator closed forms;
  * term numbering / counts -- brute-force structural counts;
  * Lanczos  -- an independent numpy Lanczos on the dense matrix;
  * GIVENS   -- scipy.linalg.eigh_tridiagonal;
  * PRO      -- orthogonality invariant and agreement of the converged Ritz
                values with the dense spectrum;
  * start vector -- closed form of the weighted |0> vector.
A planted mistake in any of these (a dropped term, a transposed state pair,
the wrong sign of a step, the q^3 element without its quirk) changes a
number checked below by far more than the tolerances.
"""
import numpy as np
import pytest
import scipy.linalg as sl

from nadvibs_amd.synth import make_problem, workload
from oracle.dense import dense_hamiltonian, dense_lanczos, qpow, structural_counts
from oracle.oracle import Oracle, givens

SWEEP = [
    (1, [9], 1, 0, 0, 0), (1, [6, 5], 2, 0, 0, 0), (2, [4, 3, 5], 2, 0, 0, 0), (3, [3, 4, 2], 3, 10, 0, 30),
    (1, [6, 5], 4, 10, 5, 0), (2, [5, 1, 4], 3, 20, 0, 20), (4, [3, 3, 2, 2], 2, 0, 0, 10),
    (2, [2, 2, 2, 2, 2, 2, 2], 3, 5, 0, 0), (3, [2, 3], 4, 5, 3, 10),
]


def _prob(case, quirks, seed=3):
    ns, nf, ordr, cub, qua, offe = case
    return make_problem(ns, nf, ordr, (400, 3000), 250.0, 60.0, 40.0, cubic=cub, quartic=qua, seed=seed,
                        quirks=quirks, offdiag_even=offe)


@pytest.mark.parametrize("quirks", [3, 0, 1, 2])
@pytest.mark.parametrize("case", SWEEP, ids=[f"s{i}" for i in range(len(SWEEP))])
def test_hv_matches_dense(case, quirks):
    p = _prob(case, quirks)
    H = dense_hamiltonian(p)
    o = Oracle(p)
    for seed in range(2):
        x = np.random.default_rng(seed).standard_normal(p.n)
        ref = H @ x
        assert np.max(np.abs(o.hv(x) - ref)) <= 1e-14 * np.max(np.abs(ref)) * 10


@pytest.mark.parametrize("case", SWEEP, ids=[f"s{i}" for i in range(len(SWEEP))])
def test_term_counts_match_structure(case):
    p = _prob(case, 3)
    np.testing.assert_array_equal(Oracle(p).term_counts(), structural_counts(p))


def test_q3_quirk_is_one_third():
    """Only the <m|q^3|m+-1> elements change between the conventions, by 3x."""
    p_ref = make_problem(1, [8], 3, (1000, 1000), 0.0, 0.0, 0.0, cubic=50.0, seed=1, quirks=1)
    p_ex = make_problem(1, [8], 3, (1000, 1000), 0.0, 0.0, 0.0, cubic=50.0, seed=1, quirks=0)
    Hr = dense_hamiltonian(p_ref) - np.diag(np.diag(dense_hamiltonian(p_ref)))
    He = dense_hamiltonian(p_ex) - np.diag(np.diag(dense_hamiltonian(p_ex)))
    x = np.random.default_rng(0).standard_normal(8)
    E = np.diag(dense_hamiltonian(p_ref))
    ar, ae = Oracle(p_ref).hv(x) - E * x, Oracle(p_ex).hv(x) - E * x
    np.testing.assert_allclose(ar, Hr @ x, atol=1e-15)
    np.testing.assert_allclose(ae, He @ x, atol=1e-15)
    d1 = np.diag(He, 1)
    d3 = np.diag(He, 3)
    np.testing.assert_allclose(np.diag(Hr, 1), d1 / 3.0, rtol=1e-14)
    np.testing.assert_allclose(np.diag(Hr, 3), d3, rtol=1e-14)


def test_qpow_is_ladder_algebra():
    """The dense factors themselves: <m|q^2|m> = m + 1/2, <m|q|m+1> = sqrt((m+1)/2),
    <m|q^4|m> = 3(2m^2+2m+1)/4 (textbook values)."""
    m = np.arange(10)
    np.testing.assert_allclose(np.diag(qpow(10, 2)), m + 0.5, rtol=1e-14)
    np.testing.assert_allclose(np.diag(qpow(10, 1), 1), np.sqrt((m[:-1] + 1) / 2), rtol=1e-14)
    np.testing.assert_allclose(np.diag(qpow(10, 4)), 3 * (2 * m * m + 2 * m + 1) / 4, rtol=1e-13)
    # the truncated q^2 is not (PqP)^2 at the edge
    q = qpow(10, 1)
    assert abs((q @ q)[9, 9] - 4.5) < 1e-12 and abs(qpow(10, 2)[9, 9] - 9.5) < 1e-12


def test_zero_coupling_closed_form():
    p = make_problem(3, [4, 5, 3], 2, (500, 2500), 0.0, 0.0, 0.0, gap=1500.0, seed=4)
    o = Oracle(p)
    x = np.random.default_rng(1).standard_normal(p.n)
    grids = np.meshgrid(*[np.arange(v) for v in p.nfunc], indexing="ij")
    E = sum((g.reshape(-1) + 0.5) * w for g, w in zip(grids, p.aomega))
    ref = np.concatenate([(E + p.concoef[s * (s + 1) // 2 + s]) * x[s * p.dimen:(s + 1) * p.dimen]
                          for s in range(3)])
    np.testing.assert_allclose(o.hv(x), ref, rtol=1e-14)


def test_displaced_oscillator_ground_state():
    """One state, linear couplings only: E0 = sum w/2 - sum kappa^2/(2w)."""
    p = make_problem(1, [40, 40], 1, (800, 2000), 150.0, 0.0, 0.0, seed=9)
    kappa = np.array([p.potterms[0, 0, k] for k in range(2)]) / np.sqrt(p.aomega)
    exact = np.sum(p.aomega) / 2 - np.sum(kappa ** 2 / (2 * p.aomega))
    o = Oracle(p)
    r = o.lanczos(120)
    ev, _, bj, _ = givens(r["alpha"], r["beta"])
    assert abs(ev[0] - exact) < 1e-12


def test_uncoupled_states_union_of_spectra():
    p = make_problem(2, [5, 4], 2, (600, 1800), 120.0, 30.0, 20.0, seed=5)
    p.potterms[1] = 0.0  # block (2,1): no coupling
    H = dense_hamiltonian(p)
    D = p.dimen
    w = np.sort(np.concatenate([np.linalg.eigvalsh(H[:D, :D]), np.linalg.eigvalsh(H[D:, D:])]))
    np.testing.assert_allclose(np.sort(np.linalg.eigvalsh(H)), w, rtol=1e-13)
    x = np.random.default_rng(2).standard_normal(p.n)
    np.testing.assert_allclose(Oracle(p).hv(x), H @ x, rtol=1e-12, atol=1e-16)


def test_lanczos_matches_dense_lanczos():
    p = workload("cfg1")
    o = Oracle(p)
    H = dense_hamiltonian(p)
    r = o.lanczos(40)
    a, b = dense_lanczos(H, o.start_vector(), 40)
    np.testing.assert_allclose(r["alpha"][:15], a[:15], rtol=1e-12)
    np.testing.assert_allclose(r["beta"][:15], b[:15], rtol=1e-11)


def test_start_vector_weights():
    p = make_problem(3, [3, 3], 2, (500, 900), 10.0, 0.0, 0.0, seed=2)
    p.statew = np.array([1.0, 2.0, 0.5])
    q = Oracle(p).start_vector().reshape(3, -1)
    w = p.statew / np.linalg.norm(p.statew)
    np.testing.assert_allclose(q[:, 0], w, rtol=1e-15)
    assert np.count_nonzero(q[:, 1:]) == 0


def test_converged_ritz_values_are_eigenvalues():
    p = workload("cfg1")
    o = Oracle(p)
    r = o.lanczos(300)
    ev, it, bj, _ = givens(r["alpha"], r["beta"])
    w = np.linalg.eigvalsh(dense_hamiltonian(p))
    conv = ev[bj < 1e-9]
    assert conv.size > 20
    d = np.min(np.abs(conv[:, None] - w[None, :]), axis=1)
    assert np.max(d) < 1e-9


def _tri(n, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.05, 0.5, n), np.r_[1.0, rng.uniform(1e-3, 0.05, n)]


@pytest.mark.parametrize("n", [1, 2, 3, 17, 120])
def test_givens_matches_scipy(n):
    a, b = _tri(n, n)
    ev, it, bj, sg, V = givens(a, b, vectors=True)
    w, U = sl.eigh_tridiagonal(a, b[1:n]) if n > 1 else (a.copy(), np.ones((1, 1)))
    np.testing.assert_allclose(ev, w, rtol=1e-14, atol=1e-16)
    np.testing.assert_allclose(it, U[0] ** 2, atol=1e-12)
    np.testing.assert_allclose(bj, np.abs(b[n] * U[-1]), atol=1e-12)
    np.testing.assert_allclose(np.abs(V @ V.T), np.eye(n), atol=1e-10)


def test_givens_cluster_sum():
    """Lanczos T from a long run without re-orthogonalisation has ghost pairs;
    the intensity summed over a cluster matches scipy."""
    p = workload("cfg1")
    r = Oracle(p).lanczos(300)
    a, b = r["alpha"], r["beta"]
    ev, it, _, _ = givens(a, b)
    w, U = sl.eigh_tridiagonal(a, b[1:300])
    np.testing.assert_allclose(ev, w, rtol=1e-13)
    edges = np.r_[0, np.nonzero(np.diff(w) > 1e-7)[0] + 1, len(w)]
    for lo, hi in zip(edges[:-1], edges[1:]):
        assert abs(it[lo:hi].sum() - (U[0, lo:hi] ** 2).sum()) < 1e-9


def test_pro_keeps_orthogonality_and_spectrum():
    p = workload("cfg1")
    o = Oracle(p)
    plain = o.lanczos(150, keep_basis=True)
    Qp = plain["basis"]
    loss_plain = np.max(np.abs(Qp @ Qp.T - np.eye(150)))
    for noise in (0.0, 1.0):
        r = o.lanczos(150, reorthog=1, chkorthog=30, noise=noise, keep_basis=True)
        Q = r["basis"]
        assert r["events"] > 0
        assert np.max(np.abs(Q @ Q.T - np.eye(150))) < 1e-6 < loss_plain
        ev, _, bj, _ = givens(r["alpha"], r["beta"])
        w = np.linalg.eigvalsh(dense_hamiltonian(p))
        conv = ev[bj < 1e-9]
        assert np.max(np.min(np.abs(conv[:, None] - w[None, :]), axis=1)) < 1e-9
        # no ghosts: converged values are distinct
        assert np.all(np.diff(np.sort(conv)) > 1e-9)


def test_exact_omega_mode_keeps_orthogonality():
    p = workload("cfg1")
    r = Oracle(p).lanczos(100, reorthog=2, keep_basis=True)
    Q = r["basis"]
    assert r["events"] > 0
    assert np.max(np.abs(Q @ Q.T - np.eye(100))) < 1e-6


def test_mixed_mode_needs_frequent_exact_checks():
    """The mixed mode's rule, pinned on the emulation (off-diagonal H*v in
    float, diagonal in double): with the reference's check period of 100 the
    omega model, which assumes FP64 rounding, lets orthogonality collapse
    (ghost 2-cycles, then overflow); with checks every 10 iterations and the
    mixed thresholds (MIXED_TAU, MIXED_ETA) the converged eigenvalues stay
    within 1e-3 cm^-1 of the FP64 run for 1000 iterations (half the
    dimension)."""
    from oracle.oracle import set_off32
    p = workload("cfg1")
    o = Oracle(p)
    set_off32(1)
    try:
        bad = o.lanczos(400, reorthog=1, chkorthog=100, seed=5, keep_basis=True)
    finally:
        set_off32(0)
    Q = np.nan_to_num(bad["basis"][:300])
    assert np.max(np.abs(Q @ Q.T - np.eye(300))) > 0.5
    r64 = o.lanczos(1000, reorthog=1, chkorthog=100, seed=5)
    r32 = o.lanczos(1000, reorthog=1, chkorthog=100, seed=5, mixed=True)
    assert np.all(np.isfinite(r32["beta"])) and r32["beta"].max() < 1.0
    ev64, it64, bj64, _ = givens(r64["alpha"], r64["beta"])
    ev32, it32, _, _ = givens(r32["alpha"], r32["beta"])
    conv = np.nonzero(bj64 < 1e-9)[0]
    assert conv.size > 100
    j = np.argmin(np.abs(ev32[None, :] - ev64[conv][:, None]), axis=1)
    assert np.max(np.abs(ev32[j] - ev64[conv])) * 219474.6313705 < 1e-3
    assert np.max(np.abs(it32[j] - it64[conv])) < 1e-6


@pytest.mark.parametrize("quirks", [3, 0])
@pytest.mark.parametrize("case", SWEEP[:6], ids=[f"s{i}" for i in range(6)])
def test_hv_rows_matches_dense_and_full(case, quirks):
    """The row-sampled H*v (or_hv_rows, the checker of the full-size GPU
    H*v) equals the dense brute force and the whole-vector restatement at
    every row; in the float emulation it equals or_hv's float sums."""
    from oracle.oracle import set_off32
    p = _prob(case, quirks)
    H = dense_hamiltonian(p)
    o = Oracle(p)
    x = np.random.default_rng(7).standard_normal(p.n)
    ref = H @ x
    rows = np.arange(o.dimen)
    got = o.hv_rows(x, rows).T.reshape(-1)
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))
    assert np.max(np.abs(got - o.hv(x))) <= 1e-14 * np.max(np.abs(ref))
    set_off32(True)
    try:
        full32 = o.hv(x)
        rows32 = o.hv_rows(x, rows).T.reshape(-1)
    finally:
        set_off32(False)
    assert np.max(np.abs(rows32 - full32)) <= 1e-14 * np.max(np.abs(ref))
    # a shuffled subset gives the same rows
    sub = np.random.default_rng(1).permutation(o.dimen)[: max(1, o.dimen // 3)]
    np.testing.assert_array_equal(o.hv_rows(x, sub), o.hv_rows(x, rows)[sub])


def test_parallel_hv_and_recurrence_match_serial():
    """The OpenMP restatement (the multi-core CPU baseline) against the
    serial one: H*v and the plain recurrence's coefficients."""
    p = workload("cfg1")
    o = Oracle(p)
    x = np.random.default_rng(3).standard_normal(o.n)
    y = o.hv(x)
    for threads in (1, 3):
        assert np.max(np.abs(o.hv_par(x, threads) - y)) <= 1e-13 * np.max(np.abs(y))
    part = o.hv_par(x, 2, rows=(100, 500))
    full = o.hv_par(x, 2)
    idx = np.zeros(o.dimen, bool)
    idx[100:500] = True
    mask = np.tile(idx, p.nstates)
    np.testing.assert_array_equal(part[mask], full[mask])
    assert not np.any(part[~mask])
    r = o.lanczos(40)
    st, a, b = o.lanczos_par(20, 4)
    a2, b2 = np.zeros(40), np.zeros(41)
    a2[:20], b2[:21] = a, b
    st, a2, b2 = o.lanczos_par(40, 2, st, a2, b2, i0=21)
    assert np.max(np.abs(a2 - r["alpha"]) / np.abs(r["alpha"])) < 1e-11
    assert np.max(np.abs(b2 - r["beta"]) / np.abs(r["beta"])) < 1e-11


def test_hundred_modes_pinned_and_accepted():
    """The reference's namelist holds basis(100) (Source/secondary.f90:830-833):
    a 100-mode operator (97 modes with one function) against the dense brute
    force, and the library's operator build accepts it (without a GPU its
    device setup fails after the validation)."""
    import ctypes

    from nadvibs_amd import _lib
    p = make_problem(2, [3, 2, 3] + [1] * 97, 2, (400, 3000), 250.0, 60.0, 40.0, seed=9)
    H = dense_hamiltonian(p)
    x = np.random.default_rng(1).standard_normal(p.n)
    ref = H @ x
    assert np.max(np.abs(Oracle(p).hv(x) - ref)) <= 1e-13 * np.max(np.abs(ref))
    lib = _lib.load()
    keep = [np.ascontiguousarray(p.nfunc, dtype=np.int64), np.ascontiguousarray(p.aomega),
            np.ascontiguousarray(p.concoef), np.ascontiguousarray(p.potterms)]
    desc = _lib.NvProblem(nstates=p.nstates, nmodes=p.nmodes, ordr=p.ordr, nfunc=_lib.iptr(keep[0]),
                          aomega=_lib.dptr(keep[1]), concoef=_lib.dptr(keep[2]), potterms=_lib.dptr(keep[3]),
                          ztoler=1e-20, quirks=3, nirreps=1)
    h = ctypes.c_void_p()
    rc = lib.nv_create(ctypes.byref(desc), 0, ctypes.byref(h))
    if rc == 0:
        lib.nv_destroy(h)
    else:
        msg = lib.nv_last_error(None).decode()
        assert rc == 2 and "nmodes" not in msg, (rc, msg)
    desc.nmodes = 101
    assert lib.nv_create(ctypes.byref(desc), 0, ctypes.byref(h)) == 1
    assert "nmodes" in lib.nv_last_error(None).decode()
