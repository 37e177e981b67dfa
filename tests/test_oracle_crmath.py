"""The oracle on glibc against the oracle on the correctly rounded functions of crmath.h.

libsdm_oracle_cr.so differs from libsdm_oracle.so only in pow / atan2 / atan / sin / cos /
log / log10 / acos, whose glibc versions are correctly rounded for all but ~0.1 % of
arguments (tests/test_crmath.py checks crmath.h against 120-bit mpmath).  On a packed bed
with every functor running, one step from the same state must take the same solver path on
every interaction (same stop code, same evaluation count, same pair list) and differ only in
the last bits of the few contacts where glibc misrounds an argument.  The device's FP64 form
equals the crmath oracle bit for bit (tests/test_gpu_bitexact.py), so this bounds the
device's residual against the reference's own arithmetic.
"""
import math

import numpy as np
import pytest

from sudodem_amd import configs
from sudodem_amd.scene import SuperquadricsMat


@pytest.fixture(scope="module")
def oracle_cr():
    from oracle_lib import load_oracle
    return load_oracle(cr=True)


def _bed(oracle, law):
    cfg = configs.config1(n_particles=1500, box_xy=20e-3)
    if law == "linear":
        cfg.engines = configs.linear_engines(cfg.engines[1].verletDist)
        cfg.materials = [SuperquadricsMat(label="mat1", Kn=1e5, Ks=7e4, frictionAngle=math.atan(0.5), density=2650.0),
                         SuperquadricsMat(label="wallmat", Kn=1e5, Ks=7e4, frictionAngle=0.0, density=2650.0)]
    return configs.densify(cfg, oracle, iterations=40)


# law, least share of contacts with bit-equal depth, largest relative depth / force difference
CASES = [("hertz", 0.99, 1e-11), ("linear", 0.98, 1e-9)]


@pytest.mark.parametrize("law,bit_frac,rel_max", CASES)
def test_glibc_and_crmath_oracles_same_path(oracle, oracle_cr, law, bit_frac, rel_max):
    cfg = _bed(oracle, law)
    a = cfg.open(oracle)
    b = cfg.open(oracle_cr)
    for o, c in ((oracle, a), (oracle_cr, b)):
        o.set_init_sort(c.ptr, 1)
    a.step(1)
    b.restore(a.snapshot())
    a.step(1)
    b.step(1)
    for name in ("intr.id1", "intr.id2", "diag.stop", "diag.nEval"):
        assert np.array_equal(a.get(name), b.get(name)), name
    da, db = a.get("geom.PenetrationDepth"), b.get("geom.PenetrationDepth")
    act = da > 0
    assert act.sum() > 500 and np.array_equal(act, db > 0)
    assert np.mean(da[act] == db[act]) >= bit_frac
    assert np.max(np.abs(da[act] - db[act]) / da[act]) <= rel_max
    Fa, Fb = a.get("contact.force").reshape(-1, 3)[act], b.get("contact.force").reshape(-1, 3)[act]
    assert np.max(np.max(np.abs(Fa - Fb), axis=1) / np.max(np.abs(Fa), axis=1)) <= rel_max
    # and not identical: glibc does misround some of the arguments this step evaluates
    assert np.mean(da[act] == db[act]) < 1.0 or np.any(Fa != Fb)
