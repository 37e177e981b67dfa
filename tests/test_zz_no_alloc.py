"""sdm_stress_noalloc through the C ABI against sdm_stress: bit for bit, on two calls in a
row, the second on a context with fewer interactions (the static scratch keeps the first
call's partials beyond the second call's blocks, which must not leak into its result)."""
import ctypes

import numpy as np
import pytest

from sudodem_amd import configs

pytestmark = pytest.mark.gpu


def _noalloc(lib):
    f = lib.lib.sdm_stress_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
    return f


def _both(lib, ctx, volume):
    out = np.zeros(9)
    rc = _noalloc(lib)(ctx.ptr, volume, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 0, lib.last_error(ctx.ptr)
    return out.reshape(3, 3), ctx.stress(volume)


def test_stress_noalloc_matches_stress(gpu):
    big = configs.densify(configs.config1(n_particles=3000, box_xy=24e-3), gpu, iterations=30)
    small = configs.densify(configs.config1(n_particles=400, box_xy=12e-3), gpu, iterations=30)
    g1 = big.open(gpu)
    g1.step(3)
    g2 = small.open(gpu)
    g2.step(3)
    n1, n2 = len(g1.get("intr.id1")), len(g2.get("intr.id1"))
    assert n1 > 4 * 256 and n2 < n1 // 4  # the second call covers fewer blocks
    a, b = _both(gpu, g1, 2.0e-5)
    assert np.any(a != 0) and np.array_equal(a, b)
    a, b = _both(gpu, g2, 0.0)
    assert np.any(a != 0) and np.array_equal(a, b)
    # and again on the first context after the second: the same values as before
    a2, b2 = _both(gpu, g1, 2.0e-5)
    assert np.array_equal(a2, b2)
    g1.step(2)
    a3, b3 = _both(gpu, g1, 2.0e-5)
    assert np.array_equal(a3, b3) and not np.array_equal(a3, a2)
