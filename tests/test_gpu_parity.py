"""GPU (libsdm_hip.so) against the CPU oracle on identical states.

Bars: the broadphase pair list (ids, orientation, container order) is bit-exact; bounds are
bit-exact; every contact takes the same solver path and its geometry and forces agree to
1e-10 relative, most of them bit for bit: the FP64 path's transcendentals are correctly
rounded, glibc's (this oracle's) for all but ~0.1 % of arguments.
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
    the interaction set, the contact classification and every contact's solver path agree
    exactly, and every contact's depth, normal, contact point, normal force, total force and
    torques agree to 1e-10 relative.

    The device's FP64 form equals the oracle built on correctly rounded functions bit for bit
    (tests/test_gpu_bitexact.py); glibc (this oracle's library) is correctly rounded for all
    but ~0.1 % of arguments, and the two oracles differ only in last bits on the same solver
    paths (tests/test_oracle_crmath.py): depth bit-equal for at least 99 % of the contacts,
    force for at least 95 %."""
    g = c1_small.open(gpu)
    c = c1_small.open(oracle)
    oracle.set_init_sort(c.ptr, 1)
    g.step(2)
    c.restore(g.snapshot())
    g.step(1)
    c.step(1)
    assert np.array_equal(_pairs(g), _pairs(c))
    dg, dc = g.get("geom.PenetrationDepth"), c.get("geom.PenetrationDepth")
    assert np.array_equal(dg > 0, dc > 0)  # contacts classified the same way
    act = dc > 0
    assert act.sum() > 1000
    assert np.array_equal(g.get("diag.stop"), c.get("diag.stop"))
    assert np.array_equal(g.get("diag.nEval"), c.get("diag.nEval"))
    assert np.max(_rel(dg[act], dc[act], 1e-12)) < 1e-10
    assert np.mean(dg[act] == dc[act]) >= 0.99
    fg, fc = g.get("contact.force").reshape(-1, 3), c.get("contact.force").reshape(-1, 3)
    assert np.mean(np.all(fg == fc, axis=1)[act]) >= 0.95
    for name in ("geom.normal", "geom.contactPoint", "phys.normalForce", "contact.force"):
        a = g.get(name).reshape(-1, 3)[act]
        b = c.get(name).reshape(-1, 3)[act]
        scale = np.max(np.abs(b), axis=1)
        err = np.max(np.abs(a - b), axis=1) / np.maximum(scale, 1e-300)
        assert np.all(err < 1e-10), (name, float(err.max()), int(np.sum(err >= 1e-10)))
    # torque = (x - contactPoint) x F: its natural scale is |x - cp| |F| (for near-central
    # contacts |T| is far smaller and the cross product cancels)
    pos = c.get("state.pos").reshape(-1, 3)
    F = c.get("contact.force").reshape(-1, 3)[act]
    cp = c.get("geom.contactPoint").reshape(-1, 3)[act]
    for name, ids in (("contact.torque1", "intr.id1"), ("contact.torque2", "intr.id2")):
        # positions before the step's integration are not kept; use |cp - x| at the end of
        # the step, which differs by one step's displacement
        r = np.linalg.norm(cp - pos[c.get(ids)[act]], axis=1) * np.linalg.norm(F, axis=1)
        a = g.get(name).reshape(-1, 3)[act]
        b = c.get(name).reshape(-1, 3)[act]
        err = np.max(np.abs(a - b), axis=1) / np.maximum(r, 1e-300)
        assert np.all(err < 1e-10), (name, float(err.max()))


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
