"""GPU: nv_convolve_noalloc (per-workgroup maxima in the library's static
device array, no allocation) against nv_convolve on the same inputs, on two
calls in a row, the second with fewer sticks and grid points; and its refusal
of a grid beyond NV_CONVOLVE_MAX_GRID."""
import ctypes

import numpy as np
import pytest

from nadvibs_amd import _lib
from nadvibs_amd.engine import Engine
from nadvibs_amd.synth import workload

pytestmark = pytest.mark.gpu


def _device_call(eng, x, y, grid, sigma):
    import torch
    xd, yd, gd = (torch.from_numpy(np.ascontiguousarray(v)).cuda() for v in (x, y, grid))
    env = torch.empty(grid.size, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = eng.lib.nv_convolve_noalloc(eng.h, x.size, ctypes.c_void_p(xd.data_ptr()), ctypes.c_void_p(yd.data_ptr()),
                                     grid.size, ctypes.c_void_p(gd.data_ptr()), float(sigma),
                                     ctypes.c_void_p(env.data_ptr()), None)
    assert rc == _lib.NV_OK, eng.lib.nv_last_error(eng.h).decode()
    torch.cuda.synchronize()   # the context's stream is among the device's streams
    return env.cpu().numpy()


def test_noalloc_convolve_matches_allocating_one_twice():
    rng = np.random.default_rng(8)
    with Engine(workload("cfg1")) as eng:
        for n, ng in ((3000, 40000), (700, 4096)):
            x = rng.uniform(0.0, 6000.0, n)
            y = rng.uniform(0.0, 1.0, n)
            grid = np.linspace(0.0, 6000.0, ng)
            ref = eng.convolve(x, y, grid, 15.0)
            got = _device_call(eng, x, y, grid, 15.0)
            print(f"n={n} ng={ng}: max |diff| {np.max(np.abs(got - ref)):.3e}", flush=True)
            np.testing.assert_array_equal(got, ref)


def test_noalloc_convolve_refuses_a_grid_beyond_its_scratch():
    import torch
    with Engine(workload("cfg1")) as eng:
        d = torch.zeros(4, dtype=torch.float64, device="cuda")
        p = ctypes.c_void_p(d.data_ptr())
        rc = eng.lib.nv_convolve_noalloc(eng.h, 4, p, p, 1048577, p, 15.0, p, None)
        assert rc == 1
        assert "NV_CONVOLVE_MAX_GRID" in eng.lib.nv_last_error(eng.h).decode()
