"""CPU tests of the periodic cell in the oracle (oracle/libsdm_oracle.so).

They pin the oracle's restatement of the periodic collider (insertionSortPeri,
spatialOverlapPeri, cellWrap: pkg/common/InsertionSortCollider.cpp:404-543) and of the
periodic shift of body 2 (shift2 = hSize * cellDist, InteractionLoop.cpp:100-105) to what
geometry fixes: the pair set is the brute-force set of periodically overlapping boxes (plus
the real contacts), each pair's cellDist names the image of body 2 that overlaps body 1, a
periodic cell far larger than the packing reproduces the aperiodic run bit for bit, and a
contact across the boundary equals the same contact inside the cell.
"""
import numpy as np
import pytest

from sudodem_amd import scene as S
from sudodem_amd import _native as nat


def _engines(law, verlet=1e-4, gravity=(0.0, 0.0, 0.0), damping=0.0):
    if law == 0:
        geo, ip2, lw = ([S.Ig2_Superquadrics_Superquadrics_SuperquadricsGeom()],
                        [S.Ip2_SuperquadricsMat_SuperquadricsMat_SuperquadricsPhys()], [S.SuperquadricsLaw()])
    else:
        geo, ip2, lw = ([S.Ig2_Superquadrics_Superquadrics_SuperquadricsGeom2()],
                        [S.Ip2_SuperquadricsMat2_SuperquadricsMat2_HertzMindlinPhys()], [S.SuperquadricsLaw2()])
    return [S.ForceResetter(), S.InsertionSortCollider([S.Bo1_Superquadrics_Aabb()], verletDist=verlet),
            S.InteractionLoop(geo, ip2, lw), S.NewtonIntegrator(damping=damping, gravity=gravity, isSuperquadrics=1)]


def _gas(n, seed, lo, hi, speed=0.5):
    rng = np.random.default_rng(seed)
    mat = S.SuperquadricsMat(Kn=1e5, Ks=7e4, frictionAngle=0.5, density=2650)
    bodies = []
    for _ in range(n):
        r = rng.uniform(0.3e-3, 0.6e-3, 3)
        e = rng.uniform(0.6, 1.4, 2)
        q = rng.normal(size=4)
        b = S.NewSuperquadrics_rot2(*r, *e, mat, *q)
        b.state.pos = rng.uniform(lo, hi, 3)
        b.state.vel = rng.normal(scale=speed, size=3)
        bodies.append(b)
    return mat, bodies


def _scene(lib, bodies, mats, engines, dt, cell=None):
    O = S.Omega(lib=lib)
    O.materials.extend(mats)
    O.bodies.extend(bodies)
    O.engines = engines
    O.dt = dt
    if cell is not None:
        O.periodic = True
        O.cell.setBox(*cell)
    return O


def _periodic_overlaps(bmin, bmax, L):
    """Pairs whose boxes overlap for some periodic image (closed intervals), by brute force
    over the nearest image of the box centres."""
    c = 0.5 * (bmin + bmax)
    w = bmax - bmin
    out = set()
    for i in range(len(c)):
        d = c[i + 1:] - c[i]
        d -= L * np.round(d / L)
        ok = np.all(np.abs(d) <= 0.5 * (w[i + 1:] + w[i]), axis=1)
        out.update((i, i + 1 + int(j)) for j in np.nonzero(ok)[0])
    return out


def test_periodic_pair_set_is_brute_force_overlap(oracle):
    """After every collider run the container holds, besides the real contacts, only pairs of
    periodically overlapping boxes, and every such pair at the latest one run later.

    QUIRK kept: insertionSortPeri starts its pass at vector index 0, not at the ring head
    loIdx (:418-421), so an inversion across the vector boundary can survive a pass; the
    crossing pair is then created at the next run.  Bodies leave the cell through every
    face (positions are not wrapped)."""
    L = np.array([6e-3, 7e-3, 8e-3])
    mat, bodies = _gas(120, 21, 0.0, 6e-3, speed=3.0)
    O = _scene(oracle, bodies, [mat], _engines(0), 1e-6, cell=L)
    O.run(1)
    crossed = lagged = 0
    late = set()
    for _ in range(60):
        i1, i2, real = O.interactions()
        real_before = {(min(a, b), max(a, b)) for a, b, r in zip(i1, i2, real) if r}
        col = O.counters()["collider_runs"]
        O.set("collider.strideActive", np.array([0], np.int32))
        O.run(1)
        assert O.counters()["collider_runs"] == col + 1
        bmin = O.get("bound.min").reshape(-1, 3)
        bmax = O.get("bound.max").reshape(-1, 3)
        i1, i2, _ = O.interactions()
        have = {(min(a, b), max(a, b)) for a, b in zip(i1, i2)}
        expect = _periodic_overlaps(bmin, bmax, L)
        assert have - expect <= real_before
        assert not ((late & expect) - have)  # a pair the last pass missed is there now
        late = expect - have
        lagged += len(late)
        pos = O.get("state.pos").reshape(-1, 3)
        crossed = max(crossed, int(((pos < 0) | (pos >= L)).any(axis=1).sum()))
    assert crossed > 10  # the test exercised bodies outside the (0,0,0) period
    assert lagged < 0.1 * len(expect) * 60


def test_periodic_celldist_is_overlapping_image(oracle):
    """cellDist of a pair is the period of body 2 relative to body 1: body 2's box moved by
    cellDist * L overlaps body 1's box (spatialOverlapPeri, :536)."""
    L = np.array([6e-3, 6e-3, 6e-3])
    mat, bodies = _gas(150, 22, -3e-3, 9e-3, speed=0.0)  # positions over several periods
    O = _scene(oracle, bodies, [mat], _engines(0), 1e-6, cell=L)
    O.run(1)
    i1, i2, _ = O.interactions()
    cd = O.get("intr.cellDist").reshape(-1, 3)
    bmin = O.get("bound.min").reshape(-1, 3)
    bmax = O.get("bound.max").reshape(-1, 3)
    assert len(i1) > 50 and (cd != 0).any()
    s = cd * L
    assert np.all(bmin[i1] <= bmax[i2] + s) and np.all(bmin[i2] + s <= bmax[i1])


def test_large_periodic_cell_equals_aperiodic_run(oracle):
    """A periodic cell much larger than a packing that never reaches its faces changes
    nothing: the ring sort degenerates into the plain insertion sort, shift2 is zero, and the
    whole state after 60 steps is bitwise that of the aperiodic scene."""
    fields = ("state.pos", "state.vel", "state.angVel", "state.ori", "intr.id1", "intr.id2",
              "geom.PenetrationDepth", "geom.normal", "phys.shearForce", "phys.normalForce")
    out = []
    for periodic in (False, True):
        mat, bodies = _gas(120, 23, 0.02, 0.025, speed=0.3)
        O = _scene(oracle, bodies, [mat], _engines(0, verlet=5e-5, damping=0.2), 1e-6,
                   cell=(0.05, 0.05, 0.05) if periodic else None)
        contacts = 0
        for _ in range(6):
            O.run(10)
            contacts = max(contacts, int((O.get("geom.PenetrationDepth") > 0).sum()))
        out.append({k: O.get(k) for k in fields})
        assert O.counters()["collider_runs"] > 2 and contacts > 5
    for k in fields:
        assert np.array_equal(out[0][k], out[1][k]), k


def _two_bodies(mat, x1, x2):
    a = S.NewSuperquadrics_rot2(0.6e-3, 0.5e-3, 0.45e-3, 0.8, 1.2, mat, 0.9, 0.2, -0.3, 0.1)
    b = S.NewSuperquadrics_rot2(0.55e-3, 0.6e-3, 0.5e-3, 1.3, 0.7, mat, 0.7, -0.4, 0.2, 0.5)
    a.state.pos = np.array(x1, float)
    b.state.pos = np.array(x2, float)
    b.state.vel = np.array([0.02, -0.01, 0.015])
    return [a, b]


def test_contact_across_the_boundary_equals_interior_contact(oracle):
    """Linear law (Geom with shift2, SuperquadricsLaw with the shifted box test): a pair that
    touches across the x face gives, in its first step, the contact of the same pair placed
    inside the cell, up to the rounding of the different coordinates (the solver stops at a
    loose tolerance, |f.c| > 0.95, so its last iterate carries that rounding at ~1e-8).

    QUIRK kept: SuperquadricsLaw applies body 2's torque about its unshifted position
    (Superquadrics.cpp:1353-1356), so that torque differs by -shift2 x F."""
    L = (8e-3, 8e-3, 8e-3)
    res = []
    for x1, x2 in (([0.3e-3, 4e-3, 4e-3], [7.4e-3, 4.2e-3, 3.9e-3]),
                   ([4.3e-3, 4e-3, 4e-3], [3.4e-3, 4.2e-3, 3.9e-3])):
        mat = S.SuperquadricsMat(Kn=1e5, Ks=7e4, frictionAngle=0.5, density=2650)
        O = _scene(oracle, _two_bodies(mat, x1, x2), [mat], _engines(0), 1e-6, cell=L)
        O.run(1)
        assert len(O.interactions()[0]) == 1
        r = {n: O.get(n).reshape(-1) for n in ("geom.PenetrationDepth", "geom.normal", "phys.normalForce",
                                               "intr.cellDist", "contact.force", "contact.torque1",
                                               "contact.torque2")}
        r["id1"] = int(O.get("intr.id1")[0])
        res.append(r)
    cross, inner = res
    assert cross["id1"] == inner["id1"]
    assert inner["geom.PenetrationDepth"][0] > 1e-5
    assert inner["intr.cellDist"].tolist() == [0, 0, 0] and abs(cross["intr.cellDist"]).sum() == 1
    assert cross["geom.PenetrationDepth"][0] == pytest.approx(inner["geom.PenetrationDepth"][0], rel=1e-9)
    assert np.allclose(cross["geom.normal"], inner["geom.normal"], rtol=0, atol=1e-7)
    F = inner["contact.force"]
    tol = 1e-7 * np.abs(F).max()
    for n in ("phys.normalForce", "contact.force"):
        assert np.allclose(cross[n], inner[n], rtol=0, atol=tol), n
    arm = tol * 8e-3
    assert np.allclose(cross["contact.torque1"], inner["contact.torque1"], rtol=0, atol=arm)
    shift2 = cross["intr.cellDist"] * np.array(L)  # the image of body id2
    assert np.allclose(cross["contact.torque2"], inner["contact.torque2"] - np.cross(shift2, F), rtol=0, atol=arm)


def test_hertz_ignores_the_periodic_shift(oracle):
    """QUIRK kept: Geom2 evaluates the unshifted position of body 2 (Superquadrics_Ig2.cpp:2328)
    and SuperquadricsLaw2 tests the unshifted boxes (Superquadrics.cpp:1381), so a Hertz
    contact across the boundary never forms; the pair stays a potential interaction."""
    L = (8e-3, 8e-3, 8e-3)
    mat = S.SuperquadricsMat2(G=2.7e7, nu=0.3, frictionAngle=0.5, density=2650)
    O = _scene(oracle, _two_bodies(mat, [0.3e-3, 4e-3, 4e-3], [7.4e-3, 4.2e-3, 3.9e-3]), [mat],
               _engines(1), 1e-6, cell=L)
    O.run(5)
    i1, i2, real = O.interactions()
    assert len(i1) == 1 and not real[0]
    assert O.get("intr.cellDist").tolist() == [1, 0, 0] or O.get("intr.cellDist").tolist() == [-1, 0, 0]
    assert np.all(O.get("forces.force") == 0)


def test_periodic_rejects_walls_and_bad_cells(oracle):
    mat = S.SuperquadricsMat(Kn=1e5, Ks=7e4, frictionAngle=0.5, density=2650)
    w = S.utils.wall(0.0, axis=2, sense=1, material=mat)
    O = _scene(oracle, [w] + _two_bodies(mat, [1e-3, 1e-3, 1e-3], [3e-3, 1e-3, 1e-3]), [mat],
               [S.ForceResetter(), S.InsertionSortCollider([S.Bo1_Superquadrics_Aabb(), S.Bo1_Wall_Aabb()],
                                                           verletDist=1e-4),
                S.InteractionLoop([S.Ig2_Wall_Superquadrics_SuperquadricsGeom(),
                                   S.Ig2_Superquadrics_Superquadrics_SuperquadricsGeom()],
                                  [S.Ip2_SuperquadricsMat_SuperquadricsMat_SuperquadricsPhys()], [S.SuperquadricsLaw()]),
                S.NewtonIntegrator(isSuperquadrics=1)], 1e-6, cell=(8e-3, 8e-3, 8e-3))
    with pytest.raises(RuntimeError, match="periodic"):
        O.run(1)
    p = S.engines_to_params(_engines(0), 1e-6, cell=(1e-3, 0.0, 1e-3))
    with pytest.raises(RuntimeError):
        nat.Context(oracle, p, 2, [mat._native()])
