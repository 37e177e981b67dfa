"""Periodic cell on the GPU (libsdm_hip.so) against the oracle and against brute force.

Bars:
 * first collider run: the interaction list (ids, orientation, container order), every
   pair's cellDist and the three bound rings (handles, periods, loIdx) are bit-exact against
   the oracle's sorted-ring sweep;
 * every later run: the whole ordered interaction list (ids, orientation, container order,
   cellDist) and the rings after insertionSortPeri equal the oracle's, run after run, over
   more than 30 collider runs (against the oracle built on the correctly rounded functions,
   whose trajectories the FP64 form follows bit for bit), and from identical states against
   the glibc oracle too; the pairs the reference creates one run late (its walk starts at
   vector index 0) are created late here as well;
 * one step from identical states (linear law, FP64): the same solver path on every contact,
   depth, normal and force within 1e-8 on every contact and 1e-10 on 99.5 % of them
   (_contact_agreement).
"""
import numpy as np
import pytest

from sudodem_amd import configs

pytestmark = pytest.mark.gpu


def _pairs(ctx):
    return np.stack([ctx.get("intr.id1"), ctx.get("intr.id2")], axis=1)


def _contact_agreement(g, c, what, bit_frac=0.0, sel=None):
    """Per-contact depth / normal / force agreement between the device and the oracle, on
    every contact.

    The device's FP64 form equals the oracle built on correctly rounded functions bit for bit
    (tests/test_gpu_bitexact.py), and that oracle takes the same solver path as the glibc one
    on every contact with last-bit differences only (tests/test_oracle_crmath.py).  Bars:
    the same path (stop code and evaluation count) on every contact compared, every contact
    within 1e-8 relative and 99.5 % within 1e-10, a share `bit_frac` with the same depth bits,
    and contacts across a face (cellDist != 0) among them.  glibc misrounds ~0.1 % of
    arguments, and the Levenberg-Marquardt forward-difference Jacobian, whose step is the last
    increment (Superquadrics_Ig2.cpp:1170), carries such a last-bit difference at O(1) weight
    into the next iterate: a few contacts per thousand end 1e-10 to 1e-9 apart.  `sel` =
    (device rows, oracle rows) of the interactions compared."""
    ig, ic = sel if sel is not None else (slice(None), slice(None))

    def G(name, k=None):
        v = g.get(name)
        return (v.reshape(-1, k) if k else v)[ig]

    def C(name, k=None):
        v = c.get(name)
        return (v.reshape(-1, k) if k else v)[ic]
    dg, dc = G("geom.PenetrationDepth"), C("geom.PenetrationDepth")
    assert np.array_equal(dg > 0, dc > 0)  # the same contacts
    act = dc > 0
    assert act.sum() > 300
    path = (G("diag.stop") == C("diag.stop")) & (G("diag.nEval") == C("diag.nEval"))
    F = C("contact.force", 3)
    Fs = np.maximum(np.max(np.abs(F), axis=1), 1e-300)
    rel = np.stack([np.abs(dg - dc) / np.maximum(dc, 1e-300),
                    np.max(np.abs(G("geom.normal", 3) - C("geom.normal", 3)), axis=1),
                    np.max(np.abs(G("contact.force", 3) - F), axis=1) / Fs], axis=1).max(axis=1)
    cross = (C("intr.cellDist", 3) != 0).any(axis=1)
    bits = np.mean(dg[act] == dc[act])
    print("%s: %d contacts (%d across a face), same path %.4f, depth bit-equal %.4f, max rel %.3g"
          % (what, act.sum(), (act & cross).sum(), path[act].mean(), bits, rel[act].max()))
    assert path[act].all()
    assert rel[act].max() <= 1e-8
    assert np.mean(rel[act] <= 1e-10) >= 0.995
    assert bits >= bit_frac
    assert (act & cross).sum() > 20


def _periodic_overlaps(bmin, bmax, L):
    c = 0.5 * (bmin + bmax)
    w = bmax - bmin
    out = set()
    for i in range(len(c)):
        d = c[i + 1:] - c[i]
        d -= L * np.round(d / L)
        ok = np.all(np.abs(d) <= 0.5 * (w[i + 1:] + w[i]), axis=1)
        out.update((i, i + 1 + int(j)) for j in np.nonzero(ok)[0])
    return out


@pytest.fixture(scope="module")
def c2_small(gpu):
    # 2,000 particles of the periodic workload's distributions, densified into contact
    cfg = configs.densify(configs.config2(n_particles=2000), gpu, iterations=80, growth=0.01, relax=20)
    print("c2_small: cell %.2f mm, solids fraction %.3f" % (cfg.cell[0] * 1e3, cfg.solids_fraction), flush=True)
    return cfg


RING = ("collider.bounds", "collider.periods", "collider.loIdx")


def _assert_lists_equal(g, c, what):
    for name in ("intr.id1", "intr.id2", "intr.cellDist") + RING:
        a, b = g.get(name), c.get(name)
        assert a.shape == b.shape and np.array_equal(a, b), "%s: %s differs" % (what, name)


def test_periodic_first_run_exact(gpu, oracle, c2_small):
    g = c2_small.open(gpu)
    c = c2_small.open(oracle)
    oracle.set_init_sort(c.ptr, 1)
    g.step(1)
    c.step(1)
    _assert_lists_equal(g, c, "first run")
    cd = g.get("intr.cellDist").reshape(-1, 3)
    assert (cd != 0).any(axis=1).sum() > 50  # pairs across the faces
    assert (g.get("collider.periods") != 0).sum() > 50  # bounds wrapped into the cell
    for name in ("bound.min", "bound.max"):
        assert np.array_equal(g.get(name), c.get(name)), name
    _contact_agreement(g, c, "first step", bit_frac=0.85)


@pytest.fixture(scope="module")
def oracle_cr():
    from oracle_lib import load_oracle
    return load_oracle(cr=True)


def test_periodic_contact_lists_exact_over_runs(gpu, oracle_cr, c2_small):
    """The whole ordered interaction list and the rings, after every collider run of a shaken
    packing (random velocities of 20 m/s; most steps force a collider run), equal the oracle's:
    insertionSortPeri's pass from vector index 0, its one-run-late pairs, its ring-head
    bookkeeping and the container's swap-with-last erase, run after run."""
    cfg = c2_small
    rng = np.random.default_rng(6)
    vel = rng.normal(scale=20.0, size=(cfg.n, 3))
    g = cfg.open(gpu)
    c = cfg.open(oracle_cr)
    oracle_cr.set_init_sort(c.ptr, 1)
    for ctx in (g, c):
        ctx.set("state.vel", vel)
        ctx.step(1)
    _assert_lists_equal(g, c, "first run")
    L = np.array(cfg.cell)
    runs = late = moved = 0
    for k in range(44):
        before = c.counters()["collider_runs"]
        if k % 4 != 3:
            for ctx in (g, c):
                ctx.set("collider.strideActive", np.array([0], np.int32))
        lo0 = c.get("collider.loIdx").copy()
        g.step(1)
        c.step(1)
        assert g.counters()["collider_runs"] == c.counters()["collider_runs"]
        if c.counters()["collider_runs"] == before:
            continue
        runs += 1
        _assert_lists_equal(g, c, "run %d (step %d)" % (runs, k + 2))
        bmin, bmax = c.get("bound.min").reshape(-1, 3), c.get("bound.max").reshape(-1, 3)
        have = {(min(a, b), max(a, b)) for a, b in zip(c.get("intr.id1"), c.get("intr.id2"))}
        late += len(_periodic_overlaps(bmin, bmax, L) - have)
        moved += int((c.get("collider.loIdx") != lo0).sum())
    print("periodic lists: %d collider runs compared, %d overlaps left for a later run, loIdx moved %d times"
          % (runs, late, moved))
    assert runs >= 30
    assert moved >= 3  # the ring heads moved
    pos = g.get("state.pos").reshape(-1, 3)
    assert ((pos < 0) | (pos >= L)).any(axis=1).sum() > 10


def test_periodic_pair_set_every_run_is_brute_force(gpu, c2_small):
    """After every collider run the container holds, besides the real contacts, only pairs of
    periodically overlapping boxes, and every such pair at the latest one run later (the
    reference's walk from vector index 0 leaves an inversion across the vector's end for the
    next run)."""
    cfg = c2_small
    rng = np.random.default_rng(3)
    g = cfg.open(gpu)
    g.set("state.vel", rng.normal(scale=20.0, size=(cfg.n, 3)))
    L = np.array(cfg.cell)
    g.step(1)
    natural = lagged = 0
    late = set()
    for k in range(30):
        i1, i2 = g.get("intr.id1"), g.get("intr.id2")
        real = (g.get("intr.hasGeom") != 0) & (g.get("intr.hasPhys") != 0)
        real_before = {(min(a, b), max(a, b)) for a, b, r in zip(i1, i2, real) if r}
        runs = g.counters()["collider_runs"]
        if k % 3 == 0:
            g.set("collider.strideActive", np.array([0], np.int32))
        g.step(1)
        if g.counters()["collider_runs"] == runs:
            continue
        natural += k % 3 != 0
        assert g.counters()["collider_runs"] == runs + 1
        bmin, bmax = g.get("bound.min").reshape(-1, 3), g.get("bound.max").reshape(-1, 3)
        i1, i2 = g.get("intr.id1"), g.get("intr.id2")
        have = {(min(a, b), max(a, b)) for a, b in zip(i1, i2)}
        assert len(have) == len(i1)  # no duplicates
        expect = _periodic_overlaps(bmin, bmax, L)
        assert have - expect <= real_before, k
        assert not ((late & expect) - have), k  # a pair the last walk missed is there now
        late = expect - have
        lagged += len(late)
        assert len(late) <= max(5, len(expect) // 50), k
    assert natural >= 5
    pos = g.get("state.pos").reshape(-1, 3)
    assert ((pos < 0) | (pos >= L)).any(axis=1).sum() > 10


def test_periodic_new_pairs_from_identical_states(gpu, oracle, c2_small):
    """From the glibc oracle's state after some motion (bodies, interactions, rings with their
    periods and heads), one forced collider run on both sides gives the same ordered list."""
    cfg = c2_small
    rng = np.random.default_rng(4)
    c = cfg.open(oracle)
    oracle.set_init_sort(c.ptr, 1)
    c.set("state.vel", rng.normal(scale=1.0, size=(cfg.n, 3)))
    c.step(8)
    g = cfg.open(gpu)
    g.restore(c.snapshot())
    _assert_lists_equal(g, c, "restored")
    before = {(min(a, b), max(a, b)) for a, b in _pairs(c)}
    for ctx in (g, c):
        ctx.set("collider.strideActive", np.array([0], np.int32))
        ctx.step(1)
    _assert_lists_equal(g, c, "after the run")
    new = [k for k in map(tuple, _pairs(c)) if (min(k), max(k)) not in before]
    assert len(new) > 20


def test_periodic_step_physics_from_identical_state(gpu, oracle, c2_small):
    cfg = c2_small
    c = cfg.open(oracle)
    oracle.set_init_sort(c.ptr, 1)
    c.step(3)
    g = cfg.open(gpu)
    g.restore(c.snapshot())
    g.step(1)
    c.step(1)
    _assert_lists_equal(g, c, "step from restored state")
    _contact_agreement(g, c, "step from restored state", bit_frac=0.9)
    xg, xc = g.get("state.pos"), c.get("state.pos")
    assert np.max(np.abs(xg - xc)) <= 1e-9 * max(1.0, np.abs(xc).max())
