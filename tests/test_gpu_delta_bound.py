"""The rounding bound of the mixed form's difference evaluation, measured on the device.

tests/delta_bound_probe (built from tests/delta_bound_probe.hip by __graft_entry__.build())
evaluates Disfun on random pairs both as a float difference from a double anchor and in double
at the same angles.  The Nelder-Mead solver takes a touching decision from the difference form
only when |c.(p2 - p1)| exceeds kDeltaErr (= 1) times the bound; these tests pin that bound:
on the shapes of the Hertz workloads (exponents U[0.6,1.4]) and on blocky ones (U[0.3,0.5]),
with the contact direction anywhere near the line of centres or within 1e-7..0.1 of a body's
principal plane, the difference form takes every evaluation, no certified touching decision
differs from the double one, and the error stays below the bound (below 0.3 of it away from
directions within 1e-11 rad of pi/2 in the support map's angle, where the double evaluation is
itself less accurate than the difference).  Exponents above 1.5 are refused (evaluated in
double)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

PROBE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "delta_bound_probe")


def _run(n, lo, hi, mode):
    if not os.path.exists(PROBE):
        raise RuntimeError("tests/delta_bound_probe is missing; build first (__graft_entry__.build())")
    r = subprocess.run([PROBE, str(n), str(lo), str(hi), str(mode)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    bands = []
    for line in r.stdout.splitlines():
        t = line.split()
        if t and t[0] == "band":
            kv = dict(zip(t[3::2], t[4::2]))
            bands.append({"lo": float(t[1]), "hi": float(t[2]), **{k: float(v) for k, v in kv.items()}})
    print(r.stdout)
    assert len(bands) == 6
    return bands


@pytest.mark.parametrize("lo,hi", [(0.6, 1.4), (0.3, 0.5)])
def test_bound_near_line_of_centres(gpu, lo, hi):
    for b in _run(1 << 21, lo, hi, 0):
        assert b["anchored"] > 10000
        assert b["in_range"] == 1.0 and b["wrong"] == 0
        assert b["max"] < 0.3, b


@pytest.mark.parametrize("lo,hi", [(0.6, 1.4), (0.3, 0.5)])
def test_bound_near_principal_planes(gpu, lo, hi):
    for b in _run(1 << 21, lo, hi, 2):
        assert b["in_range"] == 1.0 and b["small_in_range"] == 1.0 and b["wrong"] == 0
        if b["lo"] >= 1e-5:
            assert b["max"] < 0.3, b


def test_exponents_above_one_and_a_half_are_refused(gpu):
    for b in _run(1 << 18, 1.55, 1.9, 0):
        assert b["anchored"] == 0
