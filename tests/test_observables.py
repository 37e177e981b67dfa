"""Observables (getMeanCN, getStress, wall forces) in the oracle, pinned by recomputing them
from the interaction attributes and by a two-body closed form of the stress's sign and size.

Shop__meanCoordinationNumber (py/_utils.cpp:522-539), Shop::getStress (pkg/dem/Shop_02.cpp:350-377).
"""
import numpy as np
import pytest

from sudodem_amd import scene as S

import test_periodic as TP


def _state(O):
    g = lambda k: O.get(k)
    real = (g("intr.hasGeom") != 0) & (g("intr.hasPhys") != 0)
    return real, g("intr.id1"), g("intr.id2"), g("phys.normalForce").reshape(-1, 3), \
        g("phys.shearForce").reshape(-1, 3), g("state.pos").reshape(-1, 3), g("intr.cellDist").reshape(-1, 3)


def test_mean_cn_and_stress_from_attributes(oracle):
    L = np.array([6e-3, 6e-3, 6e-3])
    mat, bodies = TP._gas(150, 31, 0.0, 6e-3, speed=0.5)
    O = TP._scene(oracle, bodies, [mat], TP._engines(0), 1e-6, cell=L)
    O.run(30)
    real, i1, i2, fn, fs, pos, cd = _state(O)
    active = real & (np.linalg.norm(fn, axis=1) != 0)
    assert active.sum() > 20 and (cd[active] != 0).any()
    assert O.ctx.mean_coordination() == 2.0 * active.sum() / len(bodies)
    br = pos[i1] - pos[i2] - cd * L
    sig = np.einsum("ni,nj->ij", (fn + fs)[real], br[real]) / np.prod(L)
    got = O.ctx.stress()
    assert np.allclose(got, sig, rtol=1e-12, atol=1e-12 * np.abs(sig).max())
    assert np.allclose(O.ctx.stress(2.0), sig * np.prod(L) / 2.0, rtol=1e-12, atol=1e-12 * np.abs(sig).max())


def test_stress_two_bodies_closed_form(oracle):
    """Two round bodies pressed together along x: sigma_xx = -|F| d / V < 0 (compression), the
    other components vanish, and the force on each body is the contact force."""
    mat = S.SuperquadricsMat(Kn=1e5, Ks=7e4, frictionAngle=0.5, density=2650)
    r = 0.5e-3
    a = S.NewSuperquadrics_rot2(r, r, r, 1.0, 1.0, mat, 1, 0, 0, 0, isSphere=True)
    b = S.NewSuperquadrics_rot2(r, r, r, 1.0, 1.0, mat, 1, 0, 0, 0, isSphere=True)
    a.state.pos = np.array([2e-3, 2e-3, 2e-3])
    b.state.pos = np.array([2e-3 + 0.98e-3, 2e-3, 2e-3])
    walls = [S.utils.wall(-1.0 if s > 0 else 1.0, axis=k, sense=s, material=mat) for k in range(3) for s in (1, -1)]
    O = TP._scene(oracle, walls + [a, b], [mat], TP._engines(0), 1e-9)
    O.run(1)
    fa = O.ctx.body_forces([6, 7])
    assert fa[0, 0] < 0 and np.allclose(fa[0], -fa[1], rtol=1e-12)
    F = abs(fa[0, 0])
    assert F == pytest.approx(1e5 * (2 * r - 0.98e-3), rel=1e-6)
    sig = O.ctx.stress(1.0)
    assert sig[0, 0] == pytest.approx(-F * 0.98e-3, rel=1e-6)
    assert np.all(np.abs(sig.ravel()[1:]) <= 1e-9 * F * 1e-3)
    assert O.ctx.mean_coordination() == 2.0 * 1 / 2
