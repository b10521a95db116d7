"""The Fortran drop-in driver (driver/nadvibs.x): reads basis.in/nadvibs.in,
runs the reference's Lanczos loop through the C ABI, writes restart.log and
the output.dat eigenvalue table in the reference's formats
(Source/top.f90:1007-1112)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from nadvibs_amd.engine import Engine, read_restart_log
from nadvibs_amd.inputs import AU2EV, write_problem
from nadvibs_amd.synth import workload
from oracle.oracle import Oracle, givens

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "driver", "nadvibs.x")


def _build():
    if shutil.which("flang") is None and not os.path.exists(DRIVER):
        pytest.skip("flang is not available to build the driver")
    if shutil.which("flang") is not None:
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "driver")])


def test_driver_builds():
    """CPU: the driver compiles and links against the C ABI."""
    _build()
    assert os.path.exists(DRIVER)


def parse_table(path):
    """The columns extract.py reads (extract.py:24-49): convergence, E(eV), intensity."""
    lines = open(path).read().splitlines()
    start = stop = None
    for i, l in enumerate(lines):
        if "Root  Index  Convergence     E(cm-1)        E(eV)     Intensity" in l:
            start = i + 1
        if "All requested eigenvalues computed." in l:
            stop = i - 1
            break
    rows = [lines[start + i].split() for i in range(stop - start)]
    return np.array([[float(r[2]), float(r[3]), float(r[4]), float(r[5])] for r in rows])


def run_driver(d, env=None):
    e = dict(os.environ)
    e.update(env or {})
    out = subprocess.run([DRIVER, "-o", str(d)], cwd=d, env=e, capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout + out.stderr
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("loop", ["1", "0"])
def test_driver_cfg1_table(tmp_path, loop):
    _build()
    p = workload("cfg1")
    write_problem(tmp_path, p)
    run_driver(tmp_path, {"NV_LOOP": loop})
    tab = parse_table(tmp_path / "output.dat")
    # the driver is the library: its table is the Python binding's on the same run
    with Engine(p) as eng:
        eng.lanczos_init(p.niter)
        eng.run(1, p.niter)
        a, b, _ = eng.coefficients()
        ev_g, it_g, bj_g, _ = eng.ritz(a, b)
    keep_g = [0] + [i for i in range(1, len(ev_g)) if bj_g[i] < 1.0]
    assert len(tab) == len(keep_g)
    np.testing.assert_allclose(tab[:, 1], (ev_g[keep_g] - ev_g[0]) * 219474.6313705, atol=2e-5)
    np.testing.assert_allclose(tab[:, 3], it_g[keep_g] / max(it_g[keep_g]), atol=2e-7)
    # against the oracle: every converged root of the table is a converged root
    # of the oracle's run within 1e-4 cm^-1 (unconverged and near-degenerate
    # roots of a run without re-orthogonalisation differ with the rounding)
    r = Oracle(p).lanczos(p.niter)
    ev, it, bj, _ = givens(r["alpha"], r["beta"])
    e_ref = (ev[bj < 1e-9] - ev[0]) * 219474.6313705
    conv = tab[:, 0] < 1e-9
    assert conv.sum() > 20
    for e in tab[conv, 1]:
        assert np.min(np.abs(e_ref - e)) < 1e-4, e
    e_ref_all = (ev - ev[0]) * AU2EV
    for e in tab[conv, 2]:
        assert np.min(np.abs(e_ref_all - e)) < 2e-5
    _, n, a, b, _ = read_restart_log(tmp_path / "restart.log")
    assert n == p.niter
    np.testing.assert_allclose(a[:20], r["alpha"][:20], rtol=1e-11)
    assert os.path.getsize(tmp_path / "nadvibs.35.0") == 3 * p.n * 8


@pytest.mark.gpu
def test_driver_restart_continues_the_run(tmp_path):
    """niter=150, then restart=1 with niter=150 more, equals one 300 run."""
    _build()
    p = workload("cfg1")
    full = tmp_path / "full"
    part = tmp_path / "part"
    p.niter = 300
    write_problem(full, p)
    run_driver(full)
    p.niter = 150
    write_problem(part, p)
    run_driver(part)
    p.restart = 1
    write_problem(part, p)
    run_driver(part)
    _, n1, a1, b1, _ = read_restart_log(full / "restart.log")
    _, n2, a2, b2, _ = read_restart_log(part / "restart.log")
    assert n1 == n2 == 300
    np.testing.assert_allclose(a2, a1, rtol=1e-12)
    np.testing.assert_allclose(b2, b1, rtol=1e-12)


@pytest.mark.gpu
def test_engine_resumes_after_load_vectors(tmp_path):
    """A run restored with nv_load_vectors and continued with nv_lanczos_run
    equals one uninterrupted run (the resumed run starts with the per-call
    path, which sets the device iteration counter)."""
    p = workload("cfg1")
    with Engine(p) as eng:
        eng.lanczos_init(300)
        eng.run(1, 300)
        a1, b1, _ = eng.coefficients()
    with Engine(p) as eng:
        eng.lanczos_init(150)
        eng.run(1, 150)
        a, b, _ = eng.coefficients()
        eng.save_vectors(tmp_path / "v.bin")
    with Engine(p) as eng:
        eng.lanczos_init(300)
        eng.load_vectors(tmp_path / "v.bin", 150, a, b)
        eng.run(151, 300)
        a2, b2, _ = eng.coefficients()
    np.testing.assert_allclose(a2, a1, rtol=1e-12)
    np.testing.assert_allclose(b2, b1, rtol=1e-12)
