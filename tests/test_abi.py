"""The C ABI library loads, exports every entry point include/nadvibs.h
declares, and rejects invalid problems with a status code and a message
(no GPU needed: validation happens before any device call)."""
import ctypes
import os
import re

import numpy as np
import pytest

from nadvibs_amd import _lib
from nadvibs_amd.synth import make_problem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    text = open(os.path.join(ROOT, "include", "nadvibs.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nv_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = _lib.load()
    names = declared_entry_points()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(n for n, _, _ in _lib.SIGNATURES) == names


def _desc(p, **over):
    keep = [np.ascontiguousarray(p.nfunc, dtype=np.int64), np.ascontiguousarray(p.aomega),
            np.ascontiguousarray(p.concoef), np.ascontiguousarray(p.potterms)]
    d = dict(nstates=p.nstates, nmodes=p.nmodes, ordr=p.ordr, nfunc=_lib.iptr(keep[0]), aomega=_lib.dptr(keep[1]),
             concoef=_lib.dptr(keep[2]), potterms=_lib.dptr(keep[3]), ztoler=1e-20, quirks=3, nirreps=1)
    d.update(over)
    return _lib.NvProblem(**d), keep


@pytest.mark.parametrize("over,msg", [(dict(ordr=5), "ordr"), (dict(nstates=0), "nstates"),
                                      (dict(nstates=11), "nstates"), (dict(nmodes=0), "nmodes")])
def test_invalid_problem_is_rejected(over, msg):
    lib = _lib.load()
    p = make_problem(2, [3, 3], 2, (500, 900), 10.0, 5.0, 5.0, seed=1)
    desc, keep = _desc(p, **over)
    h = ctypes.c_void_p()
    rc = lib.nv_create(ctypes.byref(desc), 0, ctypes.byref(h))
    assert rc == 1 and not h.value
    assert msg in lib.nv_last_error(None).decode()


@pytest.mark.parametrize("ns,nfunc,msg", [(2, [2] * 31, "2^31"), (1, [2] * 101, "nmodes")])
def test_size_limits_are_refused_before_any_device_call(ns, nfunc, msg):
    """nstates * dimen from 2^31 on (here 2 x 2^31) and more than 100 modes (the
    reference's basis(100)) are refused with a message; 100 modes pass the check
    (the GPU tests run a 100-mode operator)."""
    lib = _lib.load()
    nm = len(nfunc)
    p = make_problem(ns, nfunc, 1, (500, 900), 10.0, 0.0, 0.0, seed=1)
    desc, keep = _desc(p)
    h = ctypes.c_void_p()
    rc = lib.nv_create(ctypes.byref(desc), 0, ctypes.byref(h))
    assert rc == 1 and not h.value, nm
    assert msg in lib.nv_last_error(None).decode()


def test_null_arguments():
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.nv_create(None, 0, ctypes.byref(h)) == 1
    assert lib.nv_destroy(None) == 0
    assert lib.nv_dims(None, None, None) == 1


def test_mixed_thresholds_match_between_header_and_oracle():
    """The oracle's emulation of the mixed mode uses the header's thresholds."""
    import re
    from oracle.oracle import MIXED_ETA, MIXED_TAU
    text = open(os.path.join(ROOT, "include", "nadvibs.h")).read()
    tau = float(re.search(r"#define NV_PRO_MIXED_TAU (\S+)", text).group(1))
    eta = float(re.search(r"#define NV_PRO_MIXED_ETA (\S+)", text).group(1))
    assert tau == MIXED_TAU and eta == MIXED_ETA


def test_struct_layouts_match_the_header(tmp_path):
    """ctypes mirrors of nv_problem and nv_lanczos_opts have the C layout
    (offset of every field and the size), checked with the C compiler."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "nadvibs.h"', "int main(void) {"]
    for cname, cls in (("nv_problem", _lib.NvProblem), ("nv_lanczos_opts", _lib.NvLanczosOpts)):
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.check_call([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    got = dict((" ".join(l.split()[:2]), int(l.split()[2])) for l in subprocess.check_output([str(exe)], text=True).splitlines())
    for cname, cls in (("nv_problem", _lib.NvProblem), ("nv_lanczos_opts", _lib.NvLanczosOpts)):
        assert got[f"{cname} size"] == ctypes.sizeof(cls)
        for f, _ in cls._fields_:
            assert got[f"{cname} {f}"] == getattr(cls, f).offset, (cname, f)
