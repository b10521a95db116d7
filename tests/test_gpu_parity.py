"""GPU parity of the HIP library (through the C ABI) against the oracle.

Tolerances: FP64 H*v 1e-13 relative (max norm); FP32 H*v 2e-6 relative
(off-diagonal couplings in FP32, diagonal FP64); Lanczos coefficients of the
first iterations 1e-11 relative (rounding differences grow with the
iteration count through the usual loss of orthogonality); eigenvalues of
the same tridiagonal 1e-13 relative.
"""
import os

import numpy as np
import pytest

from nadvibs_amd.engine import Engine
from nadvibs_amd.synth import make_problem, workload
from oracle.oracle import Oracle, givens

pytestmark = pytest.mark.gpu

CASES = [
    # ns, nfunc, ordr, cubic, quartic, offdiag_even, quirks
    (1, [7], 1, 0.0, 0.0, 0.0, 3),
    (2, [4, 3, 5], 2, 0.0, 0.0, 0.0, 3),
    (3, [3, 4, 2], 3, 10.0, 0.0, 30.0, 3),
    (3, [3, 4, 2], 3, 10.0, 0.0, 30.0, 0),
    (1, [6, 5], 4, 10.0, 5.0, 0.0, 3),
    (2, [5, 1, 4], 3, 20.0, 0.0, 20.0, 3),
    (4, [3, 3, 3, 2], 2, 0.0, 0.0, 10.0, 3),
    (2, [2, 2, 2, 2, 2, 2, 2], 3, 5.0, 0.0, 0.0, 3),
    (6, [3, 3, 2], 2, 0.0, 0.0, 10.0, 3),
    (9, [2, 3, 2], 3, 5.0, 0.0, 0.0, 3),
]


def _prob(case, seed=11):
    ns, nf, ordr, cub, qua, offe, quirks = case
    return make_problem(ns, nf, ordr, (400, 3000), 250.0, 60.0, 40.0, cubic=cub, quartic=qua, seed=seed,
                        quirks=quirks, offdiag_even=offe)


@pytest.fixture(params=["specialised", "generic", "skip", "steps"])
def hv_kernel(request, monkeypatch):
    """Every H*v kernel: the default (compiled at run time for the operator;
    in FP64 the shaped generic kernel where the operator allows),
    the generic one (NV_HV_KERNEL=generic; terms walked by step shape where
    the operator allows), the same with the loads of out-of-basis steps
    pointed outside the vector (the large-vector form, here forced by
    NV_HV_SKIP_MIN_BYTES=0) and the generic step-table one
    (NV_HV_KERNEL=steps)."""
    monkeypatch.delenv("NV_HV_SKIP_MIN_BYTES", raising=False)
    if request.param == "specialised":
        monkeypatch.delenv("NV_HV_KERNEL", raising=False)
    elif request.param == "skip":
        monkeypatch.setenv("NV_HV_KERNEL", "generic")
        monkeypatch.setenv("NV_HV_SKIP_MIN_BYTES", "0")
    elif request.param == "generic":
        # the plain shaped form: loads of out-of-basis steps land on other rows, times 0
        monkeypatch.setenv("NV_HV_KERNEL", "generic")
        monkeypatch.setenv("NV_HV_SKIP_MIN_BYTES", str(1 << 62))
    else:
        monkeypatch.setenv("NV_HV_KERNEL", request.param)
    return request.param


def _check_kernel(eng, kind, prec):
    info = eng.hv_kernel(prec)
    print(f"H*v kernel ({prec}): {info}", flush=True)
    if kind == "specialised":
        # the default: FP64 runs the shaped generic kernel where it can
        shaped64 = prec == "f64" and eng.prob.ordr <= 3
        assert info.startswith("generic shaped" if shaped64 else "operator-specialised"), info
    elif kind == "steps":
        assert info.startswith("generic step-table"), info
    else:
        shaped = eng.prob.ordr <= 3
        assert info.startswith("generic shaped" if shaped else "generic step-table"), info


@pytest.mark.parametrize("case", CASES, ids=[f"c{i}" for i in range(len(CASES))])
def test_hv_f64_matches_oracle(case, hv_kernel):
    p = _prob(case)
    print(f"case {case} n={p.n}", flush=True)
    orc = Oracle(p)
    with Engine(p) as eng:
        _check_kernel(eng, hv_kernel, "f64")
        np.testing.assert_array_equal(eng.term_counts(), orc.term_counts())
        np.testing.assert_array_equal(eng.term_summary(), orc.term_summary())
        for seed in range(3):
            x = np.random.default_rng(seed).standard_normal(p.n)
            y = eng.hv(x, "f64")
            ref = orc.hv(x)
            err = np.max(np.abs(y - ref)) / np.max(np.abs(ref))
            assert err < 1e-13, err


@pytest.mark.parametrize("case", CASES[:4], ids=[f"c{i}" for i in range(4)])
def test_hv_f32_matches_oracle(case, hv_kernel):
    p = _prob(case)
    orc = Oracle(p)
    with Engine(p) as eng:
        _check_kernel(eng, hv_kernel, "f32")
        x = np.random.default_rng(5).standard_normal(p.n)
        y = eng.hv(x, "f32")
        ref = orc.hv(x)
        err = np.linalg.norm(y - ref) / np.linalg.norm(ref)
        assert err < 2e-6, err


def test_hv_cfg1_and_start_vector():
    p = workload("cfg1")
    orc = Oracle(p)
    with Engine(p) as eng:
        q = eng.start_vector()
        np.testing.assert_allclose(q, orc.start_vector(), rtol=0, atol=1e-15)
        x = np.random.default_rng(3).standard_normal(p.n)
        y = eng.hv(x)
        ref = orc.hv(x)
        assert np.max(np.abs(y - ref)) / np.max(np.abs(ref)) < 1e-13


@pytest.mark.parametrize("prec,tol", [("f64", 1e-11), ("f32", 5e-6)])
def test_lanczos_coefficients_cfg1(prec, tol, hv_kernel):
    p = workload("cfg1")
    orc = Oracle(p)
    ref = orc.lanczos(40)
    with Engine(p) as eng:
        _check_kernel(eng, hv_kernel, prec)
        eng.lanczos_init(niter=40, prec=prec)
        for i in range(1, 41):  # the reference's loop, Source/main.f90:63-71
            eng.update_lanczos(i)
            eng.Hv(i)
            eng.lanczos_step(i)
        a, b, _ = eng.coefficients()
    k = 15
    np.testing.assert_allclose(a[:k], ref["alpha"][:k], rtol=tol)
    np.testing.assert_allclose(b[:k], ref["beta"][:k], rtol=tol * 10)


def _random_cases(n, seed=2024):
    """Small operators of random shape: 1-10 states, 1-5 modes of 1-6
    functions, ordr 1-3, either quirk set, with or without inter-state
    even couplings."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        nf = [int(v) for v in rng.integers(1, 7, size=int(rng.integers(1, 6)))]
        ns = int(rng.integers(1, 11))
        if ns * int(np.prod(nf)) < 8 or ns * int(np.prod(nf)) > 3000:
            continue
        out.append((ns, nf, int(rng.integers(1, 4)), float(rng.choice([0.0, 10.0])), 0.0,
                    float(rng.choice([0.0, 20.0])), int(rng.choice([0, 3]))))
    return out


@pytest.mark.parametrize("form", ["generic", "skip", "steps"])
def test_random_operators_every_generic_form(form, monkeypatch):
    """Ten random small operators through each form of the generic kernel:
    the first Lanczos coefficients in FP64 and with FP32 H*v against the
    oracle (the shaped forms cover every term shape, state count up to ten,
    modes of one function and both quirk sets)."""
    monkeypatch.setenv("NV_HV_KERNEL", "generic" if form == "skip" else form)
    if form == "skip":
        monkeypatch.setenv("NV_HV_SKIP_MIN_BYTES", "0")
    elif form == "generic":
        monkeypatch.setenv("NV_HV_SKIP_MIN_BYTES", str(1 << 62))
    else:
        monkeypatch.delenv("NV_HV_SKIP_MIN_BYTES", raising=False)
    for i, case in enumerate(_random_cases(10)):
        p = _prob(case, seed=300 + i)
        k = min(6, max(2, p.n // 4))
        print(f"random case {i}: {case} n={p.n} k={k}", flush=True)
        ref = Oracle(p).lanczos(k)
        with Engine(p) as eng:
            for prec, tol in (("f64", 1e-11), ("f32", 5e-6)):
                eng.lanczos_init(niter=k, prec=prec)
                eng.run(1, k)
                a, b, _ = eng.coefficients()
                np.testing.assert_allclose(a, ref["alpha"], rtol=tol, atol=1e-14)
                np.testing.assert_allclose(b[:k], ref["beta"][:k], rtol=tol * 10, atol=1e-14)


def test_lanczos_bf16_input_tracks_fp64():
    """BF16 copy of the H*v input (8-bit mantissa, FP32 arithmetic, FP64
    diagonal and recurrence): the first coefficients follow the FP64 oracle
    to the BF16 rounding level (2^-9 relative per element), and no further.
    The mode trades accuracy for half the FP32 vector bytes."""
    p = workload("cfg1")
    ref = Oracle(p).lanczos(40)
    with Engine(p) as eng:
        eng.lanczos_init(niter=40, prec="bf16")
        eng.run(1, 40)
        a, b, _ = eng.coefficients()
    assert np.all(np.isfinite(a)) and np.all(np.isfinite(b))
    rel = np.abs(a[:10] - ref["alpha"][:10]) / np.abs(ref["alpha"][:10])
    assert rel.max() < 1e-2
    assert rel.max() > 1e-7  # it is not running the FP32 path


@pytest.mark.parametrize("form", ["shaped", "steps", "fast-block"])
def test_generic_two_row_kernel_on_a_large_operator(form, monkeypatch):
    """4,194,304 rows in FP32, the generic H*v in its forms (the step-table
    form runs two rows per thread from this size on; the fast-block form is
    the default at this size and NV_HV_FB=0 leaves the one-pass shaped one):
    its first coefficients agree with the FP64 oracle to the FP32 level, as on
    cfg1."""
    if form == "fast-block":
        monkeypatch.delenv("NV_HV_FB", raising=False)
    else:
        monkeypatch.setenv("NV_HV_FB", "0")
    p = make_problem(2, [16, 16, 16, 16, 16, 4], 2, (400, 3000), 250.0, 60.0, 40.0, seed=13)
    assert p.n // p.nstates == 1 << 22
    ref = Oracle(p).lanczos(3)
    with Engine(p) as eng:
        eng.select_hv_kernel("steps" if form == "steps" else True)
        kind = {"shaped": "shaped", "steps": "step-table", "fast-block": "fast-block"}[form]
        assert eng.hv_kernel("f32").startswith("generic " + kind), eng.hv_kernel("f32")
        eng.lanczos_init(niter=3, prec="f32")
        eng.run(1, 3)
        a, b, _ = eng.coefficients()
    np.testing.assert_allclose(a, ref["alpha"], rtol=5e-6)
    np.testing.assert_allclose(b[1:3], ref["beta"][1:3], rtol=5e-5)


def test_ritz_matches_givens():
    p = workload("cfg1")
    ref = Oracle(p).lanczos(300)
    a, b = ref["alpha"], ref["beta"]
    with Engine(p) as eng:
        ev, it, bj, sg = eng.ritz(a, b)
    ev0, it0, bj0, sg0 = givens(a, b)
    np.testing.assert_allclose(ev, ev0, rtol=1e-13)
    # well separated roots: components agree; clusters compared through sums
    gap = np.minimum(np.r_[np.inf, np.diff(ev0)], np.r_[np.diff(ev0), np.inf])
    sep = gap > 1e-8 * np.max(np.abs(ev0))
    np.testing.assert_allclose(it[sep], it0[sep], atol=1e-10)
    np.testing.assert_allclose(bj[sep], bj0[sep], atol=1e-10)


def test_full_run_cfg1_spectrum():
    """300 iterations (the cfg1 run) on the GPU; converged Ritz values
    match the oracle's run and the dense eigenvalues."""
    from oracle.dense import dense_hamiltonian
    p = workload("cfg1")
    orc = Oracle(p)
    ref = orc.lanczos(300)
    with Engine(p) as eng:
        eng.lanczos_init(niter=300)
        eng.run(1, 300)
        a, b, _ = eng.coefficients()
        ev, it, bj, _ = eng.ritz(a, b)
    ev0, it0, bj0, _ = givens(ref["alpha"], ref["beta"])
    conv = ev[bj < 1e-9]
    conv0 = ev0[bj0 < 1e-9]
    assert len(conv) > 20
    # every converged GPU root has an oracle converged root within 1e-9 hartree
    d = np.min(np.abs(conv[:, None] - conv0[None, :]), axis=1)
    assert np.max(d) < 1e-9, np.max(d)
    w = np.linalg.eigvals(dense_hamiltonian(p)).real
    d2 = np.min(np.abs(conv[:, None] - w[None, :]), axis=1)
    assert np.max(d2) < 1e-8, np.max(d2)


def test_pro_matches_oracle_without_noise():
    """Partial re-orthogonalisation with the noise switched off.  The omega
    estimates amplify rounding differences of alpha/beta, so the range edges
    (|omega| < eps^(3/4)) may move by a few vectors; the number of events,
    the kept orthogonality and the converged spectrum must agree."""
    p = workload("cfg1")
    orc = Oracle(p)
    ref = orc.lanczos(120, reorthog=1, chkorthog=25, noise=0.0)
    with Engine(p) as eng:
        eng.lanczos_init(niter=120, reorthog=1, chkorthog=25, noise=0.0)
        eng.run(1, 120)
        a, b, om = eng.coefficients(omega=True)
        events, touched = eng.pro_stats()
        Q = np.array([eng.lanczos_vector(j) for j in range(1, 121)])
        ev, _, bj, _ = eng.ritz(a, b)
    print(f"events gpu {events} oracle {ref['events']}, touched gpu {touched} oracle {ref['touched']}")
    assert ref["events"] > 0
    assert abs(events - ref["events"]) <= 1
    assert abs(touched - ref["touched"]) <= 0.1 * ref["touched"]
    np.testing.assert_allclose(a[:40], ref["alpha"][:40], rtol=1e-9)
    G = Q @ Q.T - np.eye(Q.shape[0])
    assert np.max(np.abs(G)) < 1e-6
    ev0, _, bj0, _ = givens(ref["alpha"], ref["beta"])
    conv, conv0 = ev[bj < 1e-9], ev0[bj0 < 1e-9]
    assert conv.size > 10
    assert np.max(np.min(np.abs(conv[:, None] - conv0[None, :]), axis=1)) < 1e-9


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_captured_pro_iterations_equal_the_reference_loop(prec):
    """run() replays captured iterations, where beta's finishing sum, the
    omega recurrence and the range decision are one kernel; the reference's
    loop (update_lanczos, Hv, lanczos_step, set_omega, partial_orthog,
    Source/main.f90:63-76) launches them one by one.  Same bits: alpha, beta,
    omega and the PRO counts."""
    p = workload("cfg1")
    n = 300
    with Engine(p) as eng:
        eng.lanczos_init(niter=n, reorthog=1, prec=prec, chkorthog=50)
        eng.run(1, n)
        a1, b1, w1 = eng.coefficients(omega=True)
        s1 = eng.pro_stats()
        eng.lanczos_init(niter=n, reorthog=1, prec=prec, chkorthog=50)
        for i in range(1, n + 1):
            eng.update_lanczos(i)
            eng.Hv(i)
            eng.lanczos_step(i)
            if i > 1:
                eng.set_omega(i)
                eng.partial_orthog(i)
        a2, b2, w2 = eng.coefficients(omega=True)
        s2 = eng.pro_stats()
    print(f"PRO events/touched: {s1} {s2}", flush=True)
    assert s1 == s2 and s1[0] > 0
    assert np.array_equal(a1, a2) and np.array_equal(b1, b2) and np.array_equal(w1, w2)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_batched_graph_iterations_equal_the_reference_loop(prec):
    """Without re-orthogonalisation run() replays graphs of 16 iterations, in
    which the beta finish advances the iteration counter in place of a
    launch; runs of 1..40 iterations split into such graphs, single captured
    iterations and the per-call path give the same bits as the reference's
    loop (update_lanczos, Hv, lanczos_step)."""
    p = workload("cfg1")
    n = 75
    with Engine(p) as eng:
        eng.lanczos_init(niter=n, reorthog=0, prec=prec)
        for i in range(1, n + 1):
            eng.update_lanczos(i)
            eng.Hv(i)
            eng.lanczos_step(i)
        a0, b0, _ = eng.coefficients()
        for cuts in ([1, n], [1, 17, 18, 40, 41, n], [1, 5, 6, 38, 39, n]):
            eng.lanczos_init(niter=n, reorthog=0, prec=prec)
            for lo, hi in zip(cuts[::2], cuts[1::2]):
                eng.run(lo, hi)
            a, b, _ = eng.coefficients()
            assert np.array_equal(a, a0) and np.array_equal(b, b0), cuts


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_non_temporal_streaming_keeps_the_bits(prec, monkeypatch):
    """NV_STREAM_NT=1 (the default from 512 MB of vector) changes only the
    cache policy of the step's and normalisation's accesses: the same
    coefficients as plain accesses, to the bit."""
    p = workload("cfg1")
    n = 60
    got = []
    for nt in ("0", "1"):
        monkeypatch.setenv("NV_STREAM_NT", nt)
        with Engine(p) as eng:
            eng.lanczos_init(niter=n, reorthog=0, prec=prec)
            eng.run(1, n)
            got.append(eng.coefficients()[:2])
    assert np.array_equal(got[0][0], got[1][0]) and np.array_equal(got[0][1], got[1][1])


def test_pro_with_noise_keeps_orthogonality():
    p = workload("cfg1")
    with Engine(p) as eng:
        eng.lanczos_init(niter=150, reorthog=1, chkorthog=30, noise=1.0, seed=7)
        eng.run(1, 150)
        Q = np.array([eng.lanczos_vector(j) for j in range(1, 151)])
        events, _ = eng.pro_stats()
    assert events > 0
    assert np.max(np.abs(Q @ Q.T - np.eye(150))) < 1e-6


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("nf", [[5, 3, 3], [7, 9, 11]], ids=["N135", "N2079"])
def test_pro_odd_dimension_keeps_orthogonality(prec, nf):
    """Odd N (the FP32 basis copy is read in element pairs with a zero pad
    element per slot) at one and at two dot segments, both precisions:
    re-orthogonalisation happens and keeps the basis orthogonal.  FP64: below
    1e-6.  Mixed: FP32 rounding of H*v regrows the overlaps between the checks
    every 10 iterations; the oracle's emulation reaches 1.2e-5 on the 135-row
    problem (67 iterations, half the dimension) and 2.4e-7 on the 2079-row one,
    so the bound is 5e-5."""
    p = make_problem(3, nf, 2, (400, 3000), 250.0, 60.0, 40.0, seed=21)
    n = min(100, p.n // 2)
    with Engine(p) as eng:
        eng.lanczos_init(niter=n, reorthog=1, chkorthog=20, prec=prec, seed=3)
        eng.run(1, n)
        a, b, _ = eng.coefficients()
        Q = np.array([eng.lanczos_vector(j) for j in range(1, n + 1)])
        events, touched = eng.pro_stats()
    print(f"N {p.n} {prec}: events {events} touched {touched}")
    assert p.n % 2 == 1
    assert np.all(np.isfinite(a)) and np.all(np.isfinite(b))
    assert events > 0
    assert np.max(np.abs(Q @ Q.T - np.eye(n))) < (1e-6 if prec == "f64" else 5e-5)


@pytest.mark.parametrize("case", ["ns4", "ordr3"])
def test_mixed_thresholds_on_larger_operators(case):
    """The mixed-mode thresholds on two operators of a quarter million rows
    (four states; ordr 3 with cubic terms), 1000 iterations with partial
    re-orthogonalisation: the run stays finite and its converged roots are
    within 1e-3 cm^-1 of the FP64 run (measured 1.1e-4 and 2.7e-4)."""
    if case == "ns4":
        p = make_problem(4, [12, 10, 9, 8, 8], 2, (300, 3200), 300.0, 80.0, 40.0, seed=102)
    else:
        p = make_problem(3, [8, 8, 7, 6, 6, 5], 3, (300, 3200), 300.0, 80.0, 40.0, cubic=10.0, seed=103)
    res = {}
    with Engine(p) as eng:
        for prec in ("f64", "f32"):
            eng.lanczos_init(niter=1000, reorthog=1, chkorthog=100, prec=prec, seed=11)
            eng.run(1, 1000)
            a, b, _ = eng.coefficients()
            assert np.all(np.isfinite(a)) and np.all(np.isfinite(b)) and b.max() < 10
            res[prec] = eng.ritz(a, b)[:3]
    ev64, it64, bj64 = res["f64"]
    ev32, it32, _ = res["f32"]
    conv = np.nonzero(bj64 < 1e-9)[0]
    j = np.argmin(np.abs(ev32[None, :] - ev64[conv][:, None]), axis=1)
    dE = np.max(np.abs(ev32[j] - ev64[conv])) * 219474.6313705
    dI = np.max(np.abs(it32[j] - it64[conv]))
    print(f"{case}: max dE {dE:.2e} cm^-1, dI {dI:.1e}, {conv.size} converged")
    assert conv.size > 50
    assert dE < 1e-3 and dI < 1e-6


def test_exact_omega_mode():
    p = workload("cfg1")
    ref = Oracle(p).lanczos(80, reorthog=2)
    with Engine(p) as eng:
        eng.lanczos_init(niter=80, reorthog=2)
        eng.run(1, 80)
        a, b, _ = eng.coefficients()
        Q = np.array([eng.lanczos_vector(j) for j in range(1, 81)])
        events, _ = eng.pro_stats()
    assert events > 0 and abs(events - ref["events"]) <= 1
    assert np.max(np.abs(Q @ Q.T - np.eye(80))) < 1e-6
    np.testing.assert_allclose(a[:30], ref["alpha"][:30], rtol=1e-9)


def test_mixed_mode_long_run_against_fp64():
    """FP32 couplings + FP64 recurrence with partial re-orthogonalisation for
    1000 iterations: finite, and converged eigenvalues within 1e-3 cm^-1 of
    the FP64 oracle run; the emulated mixed oracle sees the same behaviour."""
    p = workload("cfg1")
    orc = Oracle(p)
    r64 = orc.lanczos(1000, reorthog=1, chkorthog=100, seed=5)
    rmx = orc.lanczos(1000, reorthog=1, chkorthog=100, seed=5, mixed=True)
    with Engine(p) as eng:
        eng.lanczos_init(niter=1000, reorthog=1, chkorthog=100, prec="f32", seed=5)
        eng.run(1, 1000)
        a, b, _ = eng.coefficients()
        ev, it, bj, _ = eng.ritz(a, b)
        events, touched = eng.pro_stats()
    assert np.all(np.isfinite(b)) and b.max() < 1.0
    print(f"events gpu {events} emulated oracle {rmx['events']}")
    ev64, it64, bj64, _ = givens(r64["alpha"], r64["beta"])
    conv = np.nonzero(bj64 < 1e-9)[0]
    j = np.argmin(np.abs(ev[None, :] - ev64[conv][:, None]), axis=1)
    dE = np.max(np.abs(ev[j] - ev64[conv])) * 219474.6313705
    print(f"max dE vs FP64: {dE:.2e} cm^-1")
    assert dE < 1e-3
    assert np.max(np.abs(it[j] - it64[conv])) < 1e-6


TWO_PASS = [
    # ns, nfunc, ordr, cubic, offdiag_even, quirks: operators of at most four states and three
    # or more modes, where the two-pass form applies (forced on at any size)
    (4, [4, 3, 3, 3, 2], 3, 10.0, 0.0, 3),
    (3, [3, 4, 2, 3], 3, 10.0, 30.0, 3),
    (2, [5, 2, 3], 2, 0.0, 0.0, 3),
    (4, [3, 3, 2, 2, 2, 2], 3, 5.0, 20.0, 0),
    (1, [4, 3, 3], 3, 10.0, 0.0, 3),
    (3, [1, 3, 3, 3], 3, 10.0, 0.0, 3),
    (4, [3, 1, 3, 2], 3, 10.0, 5.0, 3),
]


@pytest.mark.parametrize("variant", ["default", "mfma-skip", "valu-plain"])
@pytest.mark.parametrize("case", TWO_PASS, ids=[f"t{i}" for i in range(len(TWO_PASS))])
def test_two_pass_hv_matches_oracle(case, variant, monkeypatch):
    """The two-pass H*v (modes 0 and 1 split off) forced on small operators,
    in each form (FP32 coupling products on the VALU or on MFMA; pass 1 with
    out-of-basis loads skipped or plain): standalone H*v in both precisions
    and the first Lanczos coefficients against the oracle."""
    monkeypatch.setenv("NV_HV_TWOPASS", "1")
    monkeypatch.delenv("NV_HV_KERNEL", raising=False)
    monkeypatch.delenv("NV_HV2_MFMA", raising=False)
    monkeypatch.delenv("NV_HV2_SKIP", raising=False)
    if variant == "mfma-skip":
        monkeypatch.setenv("NV_HV2_MFMA", "1")
        monkeypatch.setenv("NV_HV2_SKIP", "1")
    elif variant == "valu-plain":
        monkeypatch.setenv("NV_HV2_MFMA", "0")
        monkeypatch.setenv("NV_HV2_SKIP", "0")
    ns, nf, ordr, cub, offe, quirks = case
    p = make_problem(ns, nf, ordr, (400, 3000), 250.0, 60.0, 40.0, cubic=cub, seed=5, quirks=quirks,
                     offdiag_even=offe)
    orc = Oracle(p)
    x = np.random.default_rng(3).standard_normal(p.n)
    ref = orc.hv(x)
    r = orc.lanczos(12)
    with Engine(p) as eng:
        for prec, tol_hv, tol_ab in (("f64", 1e-13, 1e-11), ("f32", 2e-6, 5e-6)):
            info = eng.hv_kernel(prec)
            print(f"{prec}: {info}", flush=True)
            assert info.startswith("generic two-pass"), info
            if prec == "f32":
                assert ("MFMA" in info) == (variant == "mfma-skip"), info
            err = np.max(np.abs(eng.hv(x, prec) - ref)) / np.max(np.abs(ref))
            assert err <= tol_hv, (prec, err)
            eng.lanczos_init(niter=12, reorthog=0, prec=prec)
            eng.run(1, 12)
            a, b, _ = eng.coefficients()
            assert np.max(np.abs(a - r["alpha"]) / np.abs(r["alpha"])) <= tol_ab, prec
            assert np.max(np.abs(b - r["beta"]) / np.abs(r["beta"])) <= tol_ab, prec


FAST_BLOCK = [
    # ns, nfunc, ordr, cubic, offdiag_even, quirks: operators whose trailing modes span 33..64
    # rows (the fast-block form, forced on at any size): 64, 36, 40 and 49 rows per block, slow
    # modes of one function, up to ten states, ordr 1-3, both quirk sets
    (4, [3, 4, 4, 4], 3, 10.0, 0.0, 3),
    (3, [3, 2, 6, 6], 3, 10.0, 30.0, 3),
    (2, [4, 3, 5, 8], 2, 0.0, 0.0, 3),
    (1, [5, 7, 7], 3, 10.0, 0.0, 0),
    (3, [3, 1, 4, 4, 4], 3, 10.0, 5.0, 3),
    (6, [3, 4, 4, 4], 2, 0.0, 10.0, 3),
    (2, [3, 8, 8], 1, 0.0, 0.0, 3),
    (4, [5, 2, 2, 6, 6], 3, 10.0, 10.0, 3),
    (3, [4, 1, 4, 4, 4], 3, 10.0, 5.0, 0),
    (10, [2, 5, 8], 2, 0.0, 5.0, 3),
    (4, [5, 4, 4, 4], 3, 10.0, 0.0, 3),
]


@pytest.mark.parametrize("case", FAST_BLOCK, ids=[f"b{i}" for i in range(len(FAST_BLOCK))])
def test_fast_block_hv_matches_oracle(case, monkeypatch):
    """The fast-block H*v (one wave per block of the trailing modes' rows;
    slow steps wave-uniform, fast steps through LDS) forced on small
    operators: standalone H*v in both precisions and the first Lanczos
    coefficients against the oracle."""
    monkeypatch.setenv("NV_HV_FB", "1")
    monkeypatch.delenv("NV_HV_KERNEL", raising=False)
    ns, nf, ordr, cub, offe, quirks = case
    p = make_problem(ns, nf, ordr, (400, 3000), 250.0, 60.0, 40.0, cubic=cub, seed=5, quirks=quirks,
                     offdiag_even=offe)
    print(f"case {case} n={p.n}", flush=True)
    orc = Oracle(p)
    x = np.random.default_rng(3).standard_normal(p.n)
    ref = orc.hv(x)
    r = orc.lanczos(12)
    with Engine(p) as eng:
        for prec, tol_hv, tol_ab in (("f64", 1e-13, 1e-11), ("f32", 2e-6, 5e-6)):
            info = eng.hv_kernel(prec)
            print(f"{prec}: {info}", flush=True)
            assert info.startswith("generic fast-block"), info
            err = np.max(np.abs(eng.hv(x, prec) - ref)) / np.max(np.abs(ref))
            assert err <= tol_hv, (prec, err)
            eng.lanczos_init(niter=12, reorthog=0, prec=prec)
            eng.run(1, 12)
            a, b, _ = eng.coefficients()
            assert np.max(np.abs(a - r["alpha"]) / np.abs(r["alpha"])) <= tol_ab, prec
            assert np.max(np.abs(b - r["beta"]) / np.abs(r["beta"])) <= tol_ab, prec


@pytest.mark.parametrize("kernel", ["default", "generic", "steps"])
def test_hundred_mode_operator_matches_oracle(kernel, monkeypatch):
    """100 modes (the reference's basis(100)), 96 of them with one function:
    H*v in both precisions and the first coefficients against the oracle, in
    every kernel form."""
    monkeypatch.delenv("NV_HV_KERNEL", raising=False)
    if kernel != "default":
        monkeypatch.setenv("NV_HV_KERNEL", kernel)
    # ordr 2: all 5050 quadratic monomials of 100 modes carry couplings (the checker groups the
    # terms by a linear search, as the reference does, so cubic couplings on every mode would take
    # it minutes)
    p = make_problem(2, [3, 4, 2, 3] + [1] * 96, 2, (400, 3000), 250.0, 60.0, 40.0, seed=9)
    orc = Oracle(p)
    x = np.random.default_rng(2).standard_normal(p.n)
    ref = orc.hv(x)
    r = orc.lanczos(12)
    with Engine(p) as eng:
        for prec, tol_hv, tol_ab in (("f64", 1e-13, 1e-11), ("f32", 2e-6, 5e-6)):
            err = np.max(np.abs(eng.hv(x, prec) - ref)) / np.max(np.abs(ref))
            assert err <= tol_hv, (prec, err, eng.hv_kernel(prec))
            eng.lanczos_init(niter=12, reorthog=0, prec=prec)
            eng.run(1, 12)
            a, b, _ = eng.coefficients()
            assert np.max(np.abs(a - r["alpha"]) / np.abs(r["alpha"])) <= tol_ab, prec
            assert np.max(np.abs(b - r["beta"]) / np.abs(r["beta"])) <= tol_ab, prec
