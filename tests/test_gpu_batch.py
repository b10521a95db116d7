"""GPU: batched runs (nv_batch_*) against single-problem runs and the oracle.

Every problem of a batch must reproduce its own single-problem run: the same
per-row H*v arithmetic (generic kernel) and the same recurrence, with the
dots summed over different partitions, so alpha/beta agree to 1e-11
relative over 40 iterations in FP64 (1e-8 in FP32); against the oracle the
first 20 coefficients agree to 1e-10.
"""
import numpy as np
import pytest

from nadvibs_amd.batch import Batch
from nadvibs_amd.engine import Engine
from nadvibs_amd.synth import batch_workloads, make_problem
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

SHAPES = [(1, [7, 5]), (2, [4, 3, 5]), (3, [3, 4, 2, 3]), (2, [6, 6, 5]), (3, [5, 4, 4, 3, 2]), (1, [9])]


def _probs():
    return [make_problem(ns, nf, 2, (500, 3000), 250.0, 60.0, 40.0, seed=50 + i) for i, (ns, nf) in enumerate(SHAPES)]


def _single(p, n, prec):
    with Engine(p) as eng:
        eng.select_hv_kernel(True)
        eng.lanczos_init(n, prec=prec)
        eng.run(1, n)
        return eng.coefficients()[:2]


@pytest.mark.parametrize("form", ["shaped", "steps"])
@pytest.mark.parametrize("prec,tol", [("f64", 1e-11), ("f32", 1e-8)])
def test_batch_matches_single_runs(prec, tol, form, monkeypatch):
    """Both forms of the generic H*v: terms walked by step shape, and the
    step table (NV_HV_KERNEL=steps, for batch and single runs alike)."""
    if form == "steps":
        monkeypatch.setenv("NV_HV_KERNEL", "steps")
    else:
        monkeypatch.delenv("NV_HV_KERNEL", raising=False)
    probs = _probs()
    n = 40
    with Batch(probs) as bt:
        bt.init(n, prec)
        bt.run()
        got = [bt.coefficients(i) for i in range(len(probs))]
    for p, (a, b) in zip(probs, got):
        a1, b1 = _single(p, n, prec)
        np.testing.assert_allclose(a, a1, rtol=tol, atol=1e-15)
        np.testing.assert_allclose(b, b1, rtol=tol, atol=1e-15)


@pytest.mark.parametrize("rows", ["1", "4"])
def test_batch_rows_per_thread_match_single_runs(rows, monkeypatch):
    """One and four rows per thread of the FP32 shaped batched H*v (the
    default is two) give the single-problem coefficients."""
    monkeypatch.setenv("NV_BATCH_ROWS", rows)
    monkeypatch.delenv("NV_HV_KERNEL", raising=False)
    probs = _probs()
    n = 40
    with Batch(probs) as bt:
        bt.init(n, "f32")
        bt.run()
        got = [bt.coefficients(i) for i in range(len(probs))]
    for p, (a, b) in zip(probs, got):
        a1, b1 = _single(p, n, "f32")
        np.testing.assert_allclose(a, a1, rtol=1e-8, atol=1e-15)
        np.testing.assert_allclose(b, b1, rtol=1e-8, atol=1e-15)


def test_batch_rows_setting_is_checked(monkeypatch):
    monkeypatch.setenv("NV_BATCH_ROWS", "3")
    with Batch(_probs()[:2]) as bt:
        with pytest.raises(Exception, match="NV_BATCH_ROWS"):
            bt.init(10, "f32")


def test_batch_matches_oracle_and_tuning_keeps_results():
    probs = _probs()
    n = 30
    with Batch(probs) as bt:
        bt.init(n, "f64")
        rep = bt.tune()
        assert "chosen" in rep
        bt.run(1, 10)
        bt.run(11, n)
        res = [bt.coefficients(i) for i in range(len(probs))]
        ev, it, bj, _ = bt.ritz(2, *res[2])
    for p, (a, b) in zip(probs, res):
        # a few iterations only on the tiny problems: past ~N/4 steps rounding
        # differences between any two implementations grow without bound
        k = min(20, p.n // 4)
        r = Oracle(p).lanczos(n)
        np.testing.assert_allclose(a[:k], r["alpha"][:k], rtol=1e-10)
        np.testing.assert_allclose(b[:k], r["beta"][:k], rtol=1e-10)
    assert np.all(np.diff(ev) >= 0) and abs(np.sum(it) - 1.0) < 1e-10


def test_batch_uneven_workload_sample():
    """The 64-problem uneven batch: three problems against single runs."""
    probs = batch_workloads()
    n = 12
    with Batch(probs) as bt:
        bt.init(n, "f32")
        bt.run()
        picks = [0, 17, 63]
        got = {i: bt.coefficients(i) for i in picks}
    for i in picks:
        a1, b1 = _single(probs[i], n, "f32")
        np.testing.assert_allclose(got[i][0], a1, rtol=1e-8)
        np.testing.assert_allclose(got[i][1], b1, rtol=1e-8)


def test_batch_tuner_searches_three_parameters():
    """The FP32 tuning space: rows per workgroup x rows per thread x kernel
    form (shaped or step table), every candidate timed; the chosen form gives
    the oracle's first coefficients."""
    probs = _probs()
    n = 12
    with Batch(probs) as bt:
        bt.init(n, "f32")
        rep = bt.tune()
        print(rep, flush=True)
        cands = [c for c in rep.split(":", 1)[1].split(";") if " us" in c]
        assert len(cands) >= 8, rep
        assert any("/2/" in c for c in cands) and any("steps" in c for c in cands), rep
        assert "shaped" in rep and "chosen" in rep
        bt.run(1, n)
        res = [bt.coefficients(i) for i in range(len(probs))]
    for p, (a, b) in zip(probs, res):
        k = min(8, p.n // 4)
        r = Oracle(p).lanczos(n)
        np.testing.assert_allclose(a[:k], r["alpha"][:k], rtol=5e-6)
        np.testing.assert_allclose(b[:k], r["beta"][:k], rtol=5e-6)
