"""bench.py on the CPU: its command-line contract and the pure-Python helpers
that build the result line (no GPU work)."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "--gpus 2" in r.stderr and "single" in r.stderr


def test_traffic_summaries_give_bytes_per_launch():
    for name in ("hv_f32_cfg2_traffic.txt", "hv_f32_cfg3_traffic.txt"):
        t = bench.load_traffic(os.path.join(ROOT, "profiles", name))
        assert t is not None and t > 0, name
    assert bench.load_traffic(os.path.join(ROOT, "profiles", "no_such_file.txt")) is None


def test_compare_matches_each_converged_root_to_its_nearest():
    ev64 = np.array([0.010, 0.011, 0.020, 0.030])
    it64 = np.array([0.5, 0.2, 0.2, 0.1])
    bj64 = np.array([1e-12, 1e-12, 1e-3, 1e-12])  # the third root has not converged
    # the other run: a ghost copy of the first root, a shift of 1e-9 hartree on the last
    ev = np.array([0.010, 0.010, 0.011, 0.0205, 0.030 + 1e-9])
    it = np.array([0.5, 0.0, 0.2, 0.2, 0.1 + 1e-6])
    bj = np.zeros(5)
    dE, dI, n = bench.compare(ev64, it64, bj64, ev, it, bj)
    assert n == 3
    assert abs(dE - 1e-9 * bench.AU2WAVE) < 1e-9
    assert abs(dI - 1e-6) < 1e-12


def test_compare_windows_separates_window_all_and_both_converged():
    au = bench.AU2WAVE
    ev64 = np.array([0.0, 1000.0, 5000.0, 7000.0, 2000.0]) / au
    bj64 = np.array([1e-12, 1e-12, 1e-12, 1e-12, 1e-3])
    # mixed run: the 1000 cm^-1 root off by 1e-4 cm^-1, the 7000 cm^-1 root still converging (off by 0.5)
    ev = np.array([0.0, 1000.0001, 5000.0, 7000.5, 2000.0]) / au
    bj = np.array([1e-12, 1e-12, 1e-12, 1e-4, 1e-12])
    w = bench.compare_windows(ev64, bj64, ev, bj)
    assert w["0_6000"][1] == 3 and abs(w["0_6000"][0] - 1e-4) < 1e-7
    assert w["0_6000_both"][1] == 3 and abs(w["0_6000_both"][0] - 1e-4) < 1e-7
    assert w["all"][1] == 4 and abs(w["all"][0] - 0.5) < 1e-7
    assert w["both"][1] == 3 and abs(w["both"][0] - 1e-4) < 1e-7


def test_compare_windows_leaves_roots_still_converging_out_of_the_both_figure():
    au = bench.AU2WAVE
    ev64 = np.array([0.0, 1000.0, 3000.0]) / au
    bj64 = np.full(3, 1e-12)
    # the 3000 cm^-1 root is still converging in the mixed run: off by 0.02 with residual 1e-7
    ev = np.array([0.0, 1000.00002, 3000.02]) / au
    bj = np.array([1e-12, 1e-12, 1e-7])
    w = bench.compare_windows(ev64, bj64, ev, bj)
    assert w["0_6000"][1] == 3 and abs(w["0_6000"][0] - 0.02) < 1e-7
    assert w["0_6000_both"][1] == 2 and abs(w["0_6000_both"][0] - 2e-5) < 1e-7
