"""The periodic collider's ring walk (sudodem_amd/csrc/ring.h) built for the host, against the
oracle's insertionSortPeri (InsertionSortCollider.cpp:416-463) on the oracle's own runs.

Each collider run of a shaken periodic packing (and of one drifting through the cell, which
carries loIdx into the middle of the vectors) is replayed from the rings the oracle held
before it (handles, periods, loIdx) and the bounds it computed:
 * ring_walk over the whole vector leaves exactly the oracle's rings and loIdx;
 * the device's five-phase plan (sequential phases at the start, around loIdx and at the end,
   stable sorts with crossings between them, the device's checks, extensions of sequential
   phases where a check fails) gives the same rings, coordinates, loIdx and inversions, in
   the same order, at several phase lengths;
 * the pairs the oracle appends are the overlapping new pairs in the order and orientation of
   their first inversion (axis, position inserted, rank within the insertion).
The device kernels run the same plan on one wave per axis; tests/test_gpu_periodic.py checks
them against the oracle."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from sudodem_amd import configs

HERE = os.path.dirname(os.path.abspath(__file__))
VP = ctypes.c_void_p
CAP = 1 << 20


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("ring") / "ring.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-shared", "-fPIC", "-o", out,
                    os.path.join(HERE, "ring_probe.cpp")], check=True)
    lib = ctypes.CDLL(out)
    lib.ring_whole.restype = ctypes.c_int64
    lib.ring_whole.argtypes = [VP, VP, VP, ctypes.c_int64, ctypes.c_double, VP, VP, ctypes.c_int64]
    lib.ring_phases.restype = ctypes.c_int64
    lib.ring_phases.argtypes = [VP, VP, VP, ctypes.c_int64, ctypes.c_double, VP, ctypes.c_int64, ctypes.c_int64,
                                ctypes.c_int64, VP, ctypes.c_int64, VP]
    return lib


def _run(probe, h, per, x, L, lo, phases=None):
    h, per, x = h.copy(), per.copy(), x.copy()
    lov = np.array([lo], np.int64)
    ev = np.zeros(3 * CAP, np.int64)
    st = np.zeros(2, np.int64)
    if phases is None:
        m = probe.ring_whole(h.ctypes.data, per.ctypes.data, x.ctypes.data, len(h), L, lov.ctypes.data,
                             ev.ctypes.data, CAP)
    else:
        m = probe.ring_phases(h.ctypes.data, per.ctypes.data, x.ctypes.data, len(h), L, lov.ctypes.data, *phases,
                              ev.ctypes.data, CAP, st.ctypes.data)
    assert 0 <= m <= CAP
    return h, per, x, int(lov[0]), ev[:3 * m].reshape(-1, 3), st


def _trajectory(oracle, speed, drift, steps):
    """Per collider run: the rings before it, the bounds of the run, the rings and the pairs
    after it."""
    cfg = configs.densify(configs.config2(n_particles=300, seed_extra=5), oracle, iterations=40, growth=0.01,
                          relax=10)
    ctx = cfg.open(oracle)
    oracle.set_init_sort(ctx.ptr, 1)
    rng = np.random.default_rng(11)
    ctx.set("state.vel", rng.normal(scale=speed, size=(cfg.n, 3)) + np.array([drift, 0.0, 0.0]))
    ctx.step(1)
    n2 = 2 * cfg.n
    runs = []
    for k in range(steps):
        before = {name: ctx.get(name).copy() for name in ("collider.bounds", "collider.periods", "collider.loIdx")}
        pairs0 = np.stack([ctx.get("intr.id1"), ctx.get("intr.id2")], 1)
        count = ctx.counters()["collider_runs"]
        if k % 3 != 2:
            ctx.set("collider.strideActive", np.array([0], np.int32))
        ctx.step(1)
        if ctx.counters()["collider_runs"] == count:
            continue
        runs.append(dict(before=before, bmin=ctx.get("bound.min").reshape(-1, 3), bmax=ctx.get("bound.max").reshape(-1, 3),
                         after={name: ctx.get(name).copy() for name in ("collider.bounds", "collider.periods",
                                                                         "collider.loIdx")},
                         pairs0=pairs0, pairs1=np.stack([ctx.get("intr.id1"), ctx.get("intr.id2")], 1)))
    ctx.close()
    return np.array(cfg.cell), n2, runs


@pytest.fixture(scope="module")
def trajectories(oracle):
    return [_trajectory(oracle, 20.0, 0.0, 40), _trajectory(oracle, 1.0, 30.0, 90)]


def _axis_input(run, a, n2, L):
    h = run["before"]["collider.bounds"].reshape(3, n2)[a].astype(np.int32)
    per = run["before"]["collider.periods"].reshape(3, n2)[a].astype(np.int32)
    lo = int(run["before"]["collider.loIdx"][a])
    b = np.where(h & 1, run["bmax"][h >> 1, a], run["bmin"][h >> 1, a])
    return h, per, b - (L[a] * per), lo


def test_ring_walk_is_the_oracles_pass(probe, trajectories):
    moved = 0
    for L, n2, runs in trajectories:
        assert len(runs) >= 25
        for r, run in enumerate(runs):
            for a in range(3):
                h, per, x, lo = _axis_input(run, a, n2, L)
                hw, pw, _, low, _, _ = _run(probe, h, per, x, L[a], lo)
                assert np.array_equal(hw, run["after"]["collider.bounds"].reshape(3, n2)[a]), (r, a)
                assert np.array_equal(pw, run["after"]["collider.periods"].reshape(3, n2)[a]), (r, a)
                assert low == run["after"]["collider.loIdx"][a], (r, a)
                moved += low != lo
    assert moved >= 10


@pytest.mark.parametrize("phases", [(2, 1, 8), (16, 8, 64), (64, 16, 32), (128, 64, 1024)])
def test_ring_phases_equal_the_walk(probe, trajectories, phases):
    used = np.zeros(2, np.int64)
    mid = 0
    for L, n2, runs in trajectories:
        for r, run in enumerate(runs):
            for a in range(3):
                h, per, x, lo = _axis_input(run, a, n2, L)
                hw, pw, xw, low, evw, _ = _run(probe, h, per, x, L[a], lo)
                hp, pp, xp, lop, evp, st = _run(probe, h, per, x, L[a], lo, phases)
                assert np.array_equal(hp, hw) and np.array_equal(pp, pw), (r, a)
                assert np.array_equal(xp.view(np.int64), xw.view(np.int64)), (r, a)
                assert lop == low, (r, a)
                assert np.array_equal(evp[:, 1:], evw[:, 1:]), (r, a)
                used += st
                mid += phases[0] + phases[1] < lo < n2 - phases[0] - phases[1]
    print("phases %s: %d extensions, %d sorted phases, loIdx inside the vector %d times" % (phases, *used, mid))
    assert used[1] > 0
    if phases[0] == 2:
        assert used[0] > 0 and mid > 0  # extensions and a separate zone around loIdx were exercised


def test_new_pairs_follow_the_first_inversion(probe, trajectories):
    late = 0
    for L, n2, runs in trajectories:
        for r, run in enumerate(runs):
            ev = []
            for a in range(3):
                h, per, x, lo = _axis_input(run, a, n2, L)
                ev.append(_run(probe, h, per, x, L[a], lo)[4])
            ev = np.concatenate(ev)  # axis by axis, each in the walk's order
            old = {(min(p), max(p)) for p in map(tuple, run["pairs0"].tolist())}
            new = [tuple(p) for p in run["pairs1"].tolist() if (min(p), max(p)) not in old]
            assert [tuple(p) for p in run["pairs1"][len(run["pairs1"]) - len(new):].tolist()] == new, r
            first = {}
            for _, i1, i2 in ev.tolist():
                first.setdefault((min(i1, i2), max(i1, i2)), (i1, i2))
            want = {(min(p), max(p)) for p in new}
            order = [first[k] for k in dict.fromkeys((min(i1, i2), max(i1, i2)) for _, i1, i2 in ev.tolist())
                     if k in want]
            assert order == new, r
            late += len(want - set(first))
    assert late == 0  # every new pair comes from an inversion of this run
