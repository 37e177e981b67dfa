"""The FP64 form bit for bit against the oracle built on the correctly rounded functions.

libsdm_oracle_cr.so is the oracle's restatement with pow / atan2 / atan / sin / cos / log /
log10 / acos taken from crmath.h instead of glibc (oracle/Makefile).  glibc returns the
correctly rounded value for all but ~0.1 % of arguments, so this oracle is the reference
arithmetic with glibc's misroundings removed; tests/test_oracle_pins.py shows the two oracles
take the same solver path on every contact of a Hertz bed and differ only in the last bits of
the few contacts whose arguments glibc misrounds.  Against it the device must agree on every
exported field -- bodies, interactions (order and orientation), geometry, physics, per-contact
forces and torques, solver diagnostics -- bit for bit, on every body and every contact, for the
linear law with Levenberg-Marquardt geometry (box with walls), the Hertz law with Nelder-Mead
geometry (box with walls) and the periodic cell.  Any kernel slip, whatever the contact it
touches, fails these tests; the residual against the glibc oracle (tests/test_gpu_parity.py)
is therefore glibc's own.

With SDM_DUMP_DIR set, a mismatch writes both snapshots there for inspection.
"""
import os

import numpy as np
import pytest

from sudodem_amd import configs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def oracle_cr():
    from oracle_lib import load_oracle
    return load_oracle(cr=True)


def _bits(a):
    if a.dtype.kind == "f":
        return a.view(np.int64 if a.itemsize == 8 else np.int32)
    return a


def assert_bit_equal(g, c, what, skip=()):
    """Every field of the two snapshots equal bit for bit (NaNs by their bits, signed zeros
    distinguished), except the fields named in `skip`."""
    sg, sc = g.snapshot(), c.snapshot()
    assert set(sg) == set(sc)
    bad = {}
    for name in sorted(sc):
        if name in skip:
            continue
        a, b = sg[name], sc[name]
        if a.shape != b.shape:
            bad[name] = "shape %s vs %s" % (a.shape, b.shape)
            continue
        ne = _bits(a) != _bits(b)
        if ne.any():
            bad[name] = "%d of %d differ" % (int(ne.sum()), a.size)
    if bad:
        d = os.environ.get("SDM_DUMP_DIR")
        if d:
            os.makedirs(d, exist_ok=True)
            tag = what.replace(" ", "_")
            np.savez_compressed(os.path.join(d, "bitexact_%s_gpu.npz" % tag), **{k.replace(".", "__"): v for k, v in sg.items()})
            np.savez_compressed(os.path.join(d, "bitexact_%s_cpu.npz" % tag), **{k.replace(".", "__"): v for k, v in sc.items()})
    assert not bad, "%s: %s" % (what, bad)
    return sc


def _contacts(snap):
    return int((snap["geom.PenetrationDepth"] > 0).sum())


@pytest.fixture(scope="module")
def c0_bed(gpu):
    """1,000 superellipsoids of the linear-law workload's size and shape distributions with its
    engines and materials (Levenberg-Marquardt Geom, SuperquadricsLaw, frictionless walls),
    grown from a lattice into a bed in contact so that every functor runs."""
    import math

    from sudodem_amd.scene import SuperquadricsMat
    cfg = configs.config1(n_particles=1000, box_xy=20e-3)
    cfg.engines = configs.linear_engines(cfg.engines[1].verletDist)
    cfg.materials = [SuperquadricsMat(label="mat1", Kn=1e5, Ks=7e4, frictionAngle=math.atan(0.5), density=2650.0),
                     SuperquadricsMat(label="wallmat", Kn=1e5, Ks=7e4, frictionAngle=0.0, density=2650.0)]
    return configs.densify(cfg, gpu, iterations=40)


def test_linear_lm_box_bit_exact(gpu, oracle_cr, c0_bed):
    g = c0_bed.open(gpu)
    c = c0_bed.open(oracle_cr)
    oracle_cr.set_init_sort(c.ptr, 1)
    # the bed's overlaps push it apart within ~20 steps (no confinement): every step first
    seen, it = 0, 0
    for n in [1] * 12 + [10] * 5:
        g.step(n)
        c.step(n)
        it += n
        snap = assert_bit_equal(g, c, "c0 step %d" % it)
        seen += _contacts(snap)
        if it == 1:
            assert _contacts(snap) > 300
    assert g.counters()["collider_runs"] == c.counters()["collider_runs"]
    print("c0: %d contact states compared over %d steps, %d collider runs" % (seen, it, c.counters()["collider_runs"]))


@pytest.fixture(scope="module")
def c1_bed(gpu):
    return configs.densify(configs.config1(n_particles=3000, box_xy=24e-3), gpu, iterations=60)


def test_hertz_nm_box_bit_exact(gpu, oracle_cr, c1_bed):
    g = c1_bed.open(gpu)
    c = c1_bed.open(oracle_cr)
    oracle_cr.set_init_sort(c.ptr, 1)
    g.step(1)
    c.step(1)
    snap = assert_bit_equal(g, c, "c1 first step")
    assert _contacts(snap) > 1000
    for k in range(4):
        g.step(10)
        c.step(10)
        snap = assert_bit_equal(g, c, "c1 step %d" % (1 + 10 * (k + 1)))
    print("c1: %d contacts, %d collider runs" % (_contacts(snap), c.counters()["collider_runs"]))


def test_periodic_lm_bit_exact(gpu, oracle_cr):
    cfg = configs.densify(configs.config2(n_particles=2000), gpu, iterations=80, growth=0.01, relax=20)
    g = cfg.open(gpu)
    c = cfg.open(oracle_cr)
    oracle_cr.set_init_sort(c.ptr, 1)
    g.step(1)
    c.step(1)
    snap = assert_bit_equal(g, c, "c2 first step")
    assert _contacts(snap) > 300
    assert (snap["intr.cellDist"].reshape(-1, 3) != 0).any(axis=1).sum() > 50
    # every step bit for bit, rings included, over collider runs of the reference's ring walk
    # (every other step forces one)
    for k in range(60):
        if k % 2 == 0:
            for ctx in (g, c):
                ctx.set("collider.strideActive", np.array([0], np.int32))
        g.step(1)
        c.step(1)
        snap = assert_bit_equal(g, c, "c2 step %d" % (k + 2))
    runs = c.counters()["collider_runs"]
    assert g.counters()["collider_runs"] == runs and runs >= 31
    print("c2: %d contacts, 61 steps compared, %d collider runs" % (_contacts(snap), runs))
