"""CPU: the fast-block H*v's tables, walked the way the kernel walks them.

tests/fb_emu.cpp builds each operator with the library's host code
(nadvibs_amd/csrc/nv_operator.cpp, the tables nv_create uploads) and runs the
kernel's walk on the CPU: slow monomials by phase and shape through byte
offsets and validity masks into the padded interleaved copy of x, channels
mixed per (slow monomial, fast monomial) term, fast steps through a row with
zero margins, the pure terms one by one.  It bounds-checks every read the
kernel would make and compares the result with the operator's off-diagonal
step list applied row by row (the step list itself is pinned against the
oracle's dense brute force by the GPU parity tests).
"""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from nadvibs_amd.engine import problem_struct
from nadvibs_amd.synth import make_problem, workload

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "nadvibs_amd", "csrc")

# ns, nfunc, ordr, cubic, offdiag_even, quirks: operators whose trailing modes span 33..64 rows
CASES = [
    (4, [3, 4, 4, 4], 3, 10.0, 0.0, 3),
    (3, [3, 2, 6, 6], 3, 10.0, 30.0, 3),
    (2, [4, 3, 5, 8], 2, 0.0, 0.0, 3),
    (1, [5, 7, 7], 3, 10.0, 0.0, 0),
    (3, [3, 1, 4, 4, 4], 3, 10.0, 5.0, 3),
    (6, [3, 4, 4, 4], 2, 0.0, 10.0, 3),
    (2, [3, 8, 8], 1, 0.0, 0.0, 3),
    (4, [5, 2, 2, 6, 6], 3, 10.0, 10.0, 3),
    (10, [2, 5, 8], 2, 0.0, 5.0, 3),
]


@pytest.fixture(scope="module")
def emu(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    out = str(tmp_path_factory.mktemp("fbemu") / "libfbemu.so")
    subprocess.check_call([gxx, "-O2", "-std=c++17", "-shared", "-fPIC", "-I" + os.path.join(ROOT, "include"),
                           "-I" + CSRC, os.path.join(ROOT, "tests", "fb_emu.cpp"),
                           os.path.join(CSRC, "nv_operator.cpp"), "-o", out])
    lib = ctypes.CDLL(out)
    lib.fb_emulate.restype = ctypes.c_double
    lib.fb_emulate.argtypes = [ctypes.c_void_p, ctypes.c_long] + [ctypes.POINTER(ctypes.c_int)] * 3
    return lib


def _run(lib, prob, maxblocks):
    desc, keep = problem_struct(prob)
    on, P, bad = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    err = lib.fb_emulate(ctypes.byref(desc), maxblocks, ctypes.byref(on), ctypes.byref(P), ctypes.byref(bad))
    del keep
    return err, on.value, P.value, bad.value


@pytest.mark.parametrize("case", CASES, ids=[f"c{i}" for i in range(len(CASES))])
def test_fast_block_walk_matches_step_list(emu, case):
    ns, nf, ordr, cubic, offe, quirks = case
    p = make_problem(ns, nf, ordr, (400, 3000), 250.0, 60.0, 40.0, cubic=cubic, seed=5, quirks=quirks,
                     offdiag_even=offe)
    err, on, P, bad = _run(emu, p, 1_000_000)
    assert on == 1, "the operator should take the fast-block form"
    assert 33 <= P <= 64
    assert bad == 0, f"{bad} reads out of bounds or monomials in the wrong phase"
    assert err < 1e-13, err


def test_fast_block_walk_on_the_88m_row_operator(emu):
    """The headline operator (4 states, 10 modes, ordr 3): 64 rows per block
    over modes 7..9; 300 equally spaced blocks of its 345,600."""
    p = workload("cfg3")
    err, on, P, bad = _run(emu, p, 300)
    assert on == 1 and P == 64
    assert bad == 0
    assert err < 1e-13, err


@pytest.mark.parametrize("nf", [[3, 9, 8], [2, 3, 2, 2, 4, 4], [2, 3, 3, 2, 2, 2, 2, 2, 2]])
def test_operators_without_a_fast_block(emu, nf):
    """No fast block when the fewest trailing modes that reach 33 rows span
    more than 64 (8, then 9 x 8 = 72), or when they are more than three, so
    that their monomials of degree <= 2 need more than ten channels."""
    p = make_problem(2, nf, 3, (400, 3000), 250.0, 60.0, 40.0, cubic=5.0, seed=5)
    err, on, P, bad = _run(emu, p, 10)
    assert on == 0 and bad == 0
