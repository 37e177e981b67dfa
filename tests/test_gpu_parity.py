"""GPU (libsdm_hip.so) against the CPU oracle on identical states.

Bars: the broadphase pair list (ids, orientation, container order) is bit-exact; bounds are
bit-exact; per-contact geometry and forces agree to 1e-10 relative for contacts whose solver
path is the same, and most of them bit for bit: the FP64 path's transcendentals are correctly
rounded, glibc's (the oracle's) for all but ~0.1 % of arguments, and a last-ulp difference
where glibc misrounds can flip a solver decision near a threshold (such flips are counted
and bounded).
"""
import numpy as np
import pytest

from sudodem_amd import SDM_FP64, configs

pytestmark = pytest.mark.gpu


def _pairs(ctx):
    return np.stack([ctx.get("intr.id1"), ctx.get("intr.id2")], axis=1)


@pytest.fixture(scope="module")
def c0_small():
    return configs.config0(n_particles=300)


def test_first_step_pairs_and_bounds_exact(gpu, oracle, c0_small):
    g = c0_small.open(gpu)
    c = c0_small.open(oracle)
    oracle.set_init_sort(c.ptr, 1)
    g.step(1)
    c.step(1)
    assert np.array_equal(_pairs(g), _pairs(c))
    for name in ("bound.min", "bound.max", "bound.refPos", "bound.sweepLength"):
        assert np.array_equal(g.get(name), c.get(name)), name
    assert np.array_equal(g.get("collider.bounds"), c.get("collider.bounds"))


def test_steps_state_parity(gpu, oracle, c0_small):
    g = c0_small.open(gpu)
    c = c0_small.open(oracle)
    oracle.set_init_sort(c.ptr, 1)
    for k in range(5):
        g.step(20)
        c.step(20)
        assert np.array_equal(_pairs(g), _pairs(c)), k
        pg, pc = g.get("state.pos"), c.get("state.pos")
        assert np.max(np.abs(pg - pc)) <= 1e-12 * max(1.0, np.max(np.abs(pc))), k
    assert g.counters()["collider_runs"] == c.counters()["collider_runs"]


def _rel(a, b, floor):
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


@pytest.fixture(scope="module")
def c1_small(gpu):
    # a 3,000-particle column (24x24 mm footprint) densified into a contact-carrying bed
    return configs.densify(configs.config1(n_particles=3000, box_xy=24e-3), gpu, iterations=60)


def test_hertz_nm_first_steps(gpu, oracle, c1_small):
    """Geom2 (Nelder-Mead) + Law2 on a bed with contacts, one step from identical states:
    the interaction set and the contact classification agree exactly, and per-contact depth,
    normal, contact point, normal force, total force and torques agree to 1e-10 relative.

    The device evaluates pow/log/atan2/acos/sin/cos correctly rounded, glibc (the oracle's) is
    correctly rounded for all but ~0.1 % of arguments, and the law's long double expressions
    are emulated: the depth agrees bit for bit for at least 99 % of the contacts and the force
    for at least 95 %; the rest carry last-ulp differences.  A normal is
    normalize(p1 - p2) of two surface points about `depth` apart at coordinates |x|, so its
    error is at least ulp(|x|) / depth; for the few contacts shallower than 1e-4 particle sizes
    the test allows 1e-10 or 1e3 ulp(|x|) / depth, whichever is larger."""
    g = c1_small.open(gpu)
    c = c1_small.open(oracle)
    oracle.set_init_sort(c.ptr, 1)
    g.step(2)
    c.restore(g.snapshot())
    g.step(1)
    c.step(1)
    assert np.array_equal(_pairs(g), _pairs(c))
    dg, dc = g.get("geom.PenetrationDepth"), c.get("geom.PenetrationDepth")
    act = (dc > 0) & (dg > 0)
    assert act.sum() > 1000
    # contacts classified the same way
    assert np.array_equal(dg > 0, dc > 0)
    same_path = g.get("diag.stop") == c.get("diag.stop")
    ok = act & same_path
    assert ok.sum() >= 0.99 * act.sum()
    assert np.max(_rel(dg[ok], dc[ok], 1e-12)) < 1e-10
    assert np.mean(dg[ok] == dc[ok]) >= 0.99
    # the Hertz law (curvatures, the long double stiffness expressions) on top: forces too
    fg, fc = g.get("contact.force").reshape(-1, 3), c.get("contact.force").reshape(-1, 3)
    assert np.mean(np.all(fg == fc, axis=1)[ok]) >= 0.95
    pos = c.get("state.pos").reshape(-1, 3)
    xmag = np.linalg.norm(pos[c.get("intr.id2")[ok]], axis=1)
    cond = np.maximum(1e-10, 1e3 * np.finfo(float).eps * xmag / dc[ok])
    for name in ("geom.normal", "geom.contactPoint", "phys.normalForce", "contact.force"):
        a = g.get(name).reshape(-1, 3)[ok]
        b = c.get(name).reshape(-1, 3)[ok]
        scale = np.max(np.abs(b), axis=1)
        err = np.max(np.abs(a - b), axis=1) / np.maximum(scale, 1e-300)
        tol = 1e-10 if name == "geom.contactPoint" else cond
        assert np.all(err < tol), (name, float(err.max()), int(np.sum(err >= 1e-10)))
    # torque = (x - contactPoint) x F: its natural scale is |x - cp| |F| (for near-central
    # contacts |T| is far smaller and the cross product cancels)
    F = c.get("contact.force").reshape(-1, 3)[ok]
    cp = c.get("geom.contactPoint").reshape(-1, 3)[ok]
    for name, ids in (("contact.torque1", "intr.id1"), ("contact.torque2", "intr.id2")):
        # positions before the step's integration are not kept; use |cp - x| at the end of
        # the step, which differs by one step's displacement
        r = np.linalg.norm(cp - pos[c.get(ids)[ok]], axis=1) * np.linalg.norm(F, axis=1)
        a = g.get(name).reshape(-1, 3)[ok]
        b = c.get(name).reshape(-1, 3)[ok]
        err = np.max(np.abs(a - b), axis=1) / np.maximum(r, 1e-300)
        assert np.all(err < cond), (name, float(err.max()))


def test_wall_contacts_on_every_axis(gpu, oracle):
    """Walls resting on the outermost particles of all six faces.  The collider creates the
    pairs of the walls whose planes lie on its sort axis as (particle, wall); the functor
    dispatch swaps them to (wall, particle) (InteractionLoop: swapOrder), and the force sum must
    then credit +F to the wall and -F to the particle: forces and positions follow the oracle."""
    import test_servo as TS
    g = TS.gentle_scene(gpu)[1]
    c = TS.gentle_scene(oracle)[1]
    oracle.set_init_sort(c.ptr, 1)
    for k in range(5):
        g.step(1)
        c.step(1)
        F = c.body_forces(np.arange(6))
        assert (np.abs(F) > 0).any(axis=1).all(), k  # every wall touched
        assert np.allclose(g.body_forces(np.arange(6)), F, rtol=1e-12, atol=1e-14), k
        assert np.allclose(g.get("state.pos"), c.get("state.pos"), rtol=0, atol=1e-15), k
