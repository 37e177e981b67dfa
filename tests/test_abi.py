"""The C ABI: every function declared in include/sdm.h is exported by the HIP library and,
with the sdmcpu_ prefix, by the CPU oracle (except the frozen-contact rollout, whose checker
is the forward-mode restatement in tests/frozen_ref.py).  Loading the HIP library needs no GPU."""
import ctypes
import os
import re

import pytest

from sudodem_amd import _native as nat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "sdm.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(sdm_[a-z0-9_]+)\s*\(", text))
    assert len(names) >= 15
    return sorted(names)


@pytest.mark.parametrize("name", _declared())
def test_hip_library_exports(name):
    lib = ctypes.CDLL(nat.HIP_LIB)
    assert hasattr(lib, name)


@pytest.mark.parametrize("name", [n for n in _declared() if not n.startswith("sdm_frozen_")])
def test_oracle_exports(name, oracle):
    # the frozen-contact rollout has its checker in tests/frozen_ref.py, not in the oracle
    assert hasattr(oracle.lib, "sdmcpu_" + name[len("sdm_"):])


@pytest.mark.parametrize("libname", ["hip", "oracle"])
def test_version_and_argument_errors(libname, oracle):
    """Invalid parameter sets are refused before any device work, with a message (the
    reference's InsertionSortCollider needs verletDist set when there are no spheres,
    pkg/common/InsertionSortCollider.cpp:241-252)."""
    lib = nat.load_hip() if libname == "hip" else oracle
    assert lib.version()
    p = nat.sdm_params()
    p.dt = 1e-6
    p.verlet_dist = -0.5
    assert not lib.create(ctypes.byref(p), 10, 1)
    assert b"verletDist" in lib.last_error(None)
    p.verlet_dist = 1e-4
    p.sort_axis = 3
    assert not lib.create(ctypes.byref(p), 10, 1)
    assert b"sortAxis" in lib.last_error(None)
