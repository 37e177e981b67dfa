"""The C ABI: every function declared in include/sdm.h is exported by the HIP library and,
with the sdmcpu_ prefix, by the CPU oracle (except the frozen-contact rollout, whose checker
is the forward-mode restatement in tests/frozen_ref.py, and sdm_stress_noalloc, whose checker
is sdm_stress).  Loading the HIP library needs no GPU."""
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


@pytest.mark.parametrize("name", [n for n in _declared() if not n.startswith("sdm_frozen_") and
                                  n != "sdm_stress_noalloc"])
def test_oracle_exports(name, oracle):
    # the frozen-contact rollout has its checker in tests/frozen_ref.py, not in the oracle;
    # sdm_stress_noalloc (a device-memory variant) is checked against sdm_stress
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


def check_observable_argument_errors(lib):
    """A scene of walls only: getMeanCN has no particle to divide by and reports an error
    (the reference would return inf or NaN, py/_utils.cpp:522-539); a negative interaction
    count is refused before the container is touched."""
    from sudodem_amd import scene as S
    mat = S.SuperquadricsMat(Kn=1e5, Ks=7e4, frictionAngle=0.5, density=2650)
    O = S.Omega(lib=lib)
    O.materials.append(mat)
    O.bodies.extend([S.utils.wall(-1.0 if s > 0 else 1.0, axis=a, sense=s, material=mat) for a in range(3)
                     for s in (1, -1)])
    O.engines = [S.ForceResetter(), S.InsertionSortCollider([S.Bo1_Superquadrics_Aabb(), S.Bo1_Wall_Aabb()],
                                                            verletDist=1e-4),
                 S.InteractionLoop([S.Ig2_Wall_Superquadrics_SuperquadricsGeom(),
                                    S.Ig2_Superquadrics_Superquadrics_SuperquadricsGeom()],
                                   [S.Ip2_SuperquadricsMat_SuperquadricsMat_SuperquadricsPhys()], [S.SuperquadricsLaw()]),
                 S.NewtonIntegrator(damping=0.3, gravity=(0.0, 0.0, -9.81), isSuperquadrics=1)]
    O.dt = 1e-6
    O.run(1)
    ctx = O.ctx
    with pytest.raises(nat.SdmError, match="no particles"):
        ctx.mean_coordination()
    n = len(ctx.get("intr.id1"))
    assert lib.set(ctx.ptr, b"intr.id1", None, -1) != 0
    assert b"negative" in lib.last_error(ctx.ptr)
    assert len(ctx.get("intr.id1")) == n
    O.run(1)  # the context still steps


def test_observable_argument_errors_oracle(oracle):
    check_observable_argument_errors(oracle)
