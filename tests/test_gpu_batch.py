"""Batching: several independent specimens advanced in one context give, specimen by
specimen, the result of running each specimen alone: against the device itself, against the
oracle running each specimen alone, and -- at the batching workload's full 256 specimens --
the broadphase's exact pair set per specimen (GPU)."""
import numpy as np
import pytest

from sudodem_amd import configs

pytestmark = pytest.mark.gpu


def _pairs_of(ctx, lo, hi):
    i1, i2 = ctx.get("intr.id1"), ctx.get("intr.id2")
    sel = (i1 >= lo) & (i1 < hi)
    return i1[sel] - lo, i2[sel] - lo, sel


def test_batch_equals_specimens_alone(gpu):
    rng = np.random.Generator(np.random.PCG64(20260004))
    specs = [configs.box_specimen(n, rng, kn, mu) for n, kn, mu in ((300, 6e4, 0.3), (520, 1.5e5, 0.7), (180, 1e5, 0.5))]
    vd = 0.1 * float(np.mean(np.concatenate([c.arrays["r_max"][6:] for c in specs])))
    for c in specs:
        c.engines[1].verletDist = vd
    b = configs.batch(specs)
    bc = b.open(gpu)
    alone = [c.open(gpu) for c in specs]
    for _ in range(4):
        bc.step(15)
        for c in alone:
            c.step(15)
    # no interaction joins two specimens
    spec = b.arrays["specimen"]
    assert np.array_equal(spec[bc.get("intr.id1")], spec[bc.get("intr.id2")])
    pos = bc.get("state.pos").reshape(-1, 3)
    off = 0
    for c, ctx in zip(specs, alone):
        p = ctx.get("state.pos").reshape(-1, 3)
        q = pos[off:off + c.n]
        assert np.max(np.abs(p - q)) <= 1e-12 * max(1.0, np.max(np.abs(p)))
        # same interactions (the container order may differ: erasures move the last entry)
        a1, a2, sel = _pairs_of(bc, off, off + c.n)
        s1, s2 = ctx.get("intr.id1"), ctx.get("intr.id2")
        assert sorted(zip(a1.tolist(), a2.tolist())) == sorted(zip(s1.tolist(), s2.tolist()))
        d_b = np.sort(bc.get("geom.PenetrationDepth")[sel])
        d_a = np.sort(ctx.get("geom.PenetrationDepth"))
        assert np.allclose(d_b, d_a, rtol=1e-9, atol=1e-18)
        off += c.n
    assert off == b.n


def _contact_specimens(oracle, sizes_kn_mu, seed=20260044):
    """Box specimens (the batching workload's construction) grown by the engine into contact
    (densify, evaluated by the oracle), each with its own Kn and friction coefficient."""
    rng = np.random.Generator(np.random.PCG64(seed))
    specs = []
    for n, kn, mu in sizes_kn_mu:
        c = configs.box_specimen(n, rng, kn, mu)
        c.rxyz_drawn = 1.3 * c.arrays["rxyz"]
        specs.append(configs.densify(c, oracle, iterations=60, growth=0.01, relax=20))
    vd = 0.1 * float(np.mean(np.concatenate([c.arrays["r_max"][6:] for c in specs])))
    for c in specs:
        c.engines[1].verletDist = vd
    return specs


def test_batch_against_oracle_each_specimen_alone(gpu, oracle):
    """Eight specimens of uneven sizes, own Kn in [5e4, 2e5] and friction in [0.2, 0.8],
    advanced together on the device; each specimen alone on the oracle: after every 20 steps
    the same interactions and positions within 1e-12, contact depths within 1e-8."""
    specs = _contact_specimens(oracle, ((60, 6e4, 0.3), (420, 1.5e5, 0.7), (150, 1e5, 0.5), (90, 2e5, 0.2),
                                        (300, 5e4, 0.8), (75, 1.2e5, 0.4), (200, 8e4, 0.6), (130, 1.8e5, 0.25)))
    b = configs.batch(specs)
    bc = b.open(gpu)
    alone = [c.open(oracle) for c in specs]
    for ctx in alone:
        oracle.set_init_sort(ctx.ptr, 1)
    contact_steps = 0
    for k in range(3):
        for _ in range(20):
            bc.step(1)
            contact_steps += int((bc.get("geom.PenetrationDepth") > 0).sum())
        for ctx in alone:
            ctx.step(20)
        spec = b.arrays["specimen"]
        assert np.array_equal(spec[bc.get("intr.id1")], spec[bc.get("intr.id2")])
        pos = bc.get("state.pos").reshape(-1, 3)
        off = 0
        for c, ctx in zip(specs, alone):
            p = ctx.get("state.pos").reshape(-1, 3)
            assert np.max(np.abs(p - pos[off:off + c.n])) <= 1e-12 * max(1.0, np.max(np.abs(p))), (k, c.n)
            a1, a2, sel = _pairs_of(bc, off, off + c.n)
            s1, s2 = ctx.get("intr.id1"), ctx.get("intr.id2")
            assert sorted(zip(a1.tolist(), a2.tolist())) == sorted(zip(s1.tolist(), s2.tolist())), (k, c.n)
            d_b = np.sort(bc.get("geom.PenetrationDepth")[sel])
            d_a = np.sort(ctx.get("geom.PenetrationDepth"))
            assert np.allclose(d_b, d_a, rtol=1e-8, atol=1e-18), (k, c.n)
            off += c.n
    print("batch vs oracle: %d contact-steps compared" % contact_steps)
    assert contact_steps > 10000


def _box_overlaps(bmin, bmax):
    """Keys i * n + j (i < j) of all pairs of closed boxes that overlap on the three axes,
    sorted; boxes with infinite extents (walls) are compared with every body, the finite ones
    through a k-d tree of their centres in the max norm."""
    from scipy.spatial import cKDTree
    n = len(bmin)
    fin = np.all(np.isfinite(bmin) & np.isfinite(bmax), axis=1)
    idx = np.nonzero(fin)[0]
    keys = []
    if len(idx) > 1:
        c = 0.5 * (bmin[idx] + bmax[idx])
        h = 0.5 * (bmax[idx] - bmin[idx])
        pr = cKDTree(c).query_pairs(r=2 * float(h.max()) * (1 + 1e-12), p=np.inf, output_type="ndarray")
        i, j = idx[pr[:, 0]], idx[pr[:, 1]]
        ok = np.all(bmin[i] <= bmax[j], axis=1) & np.all(bmin[j] <= bmax[i], axis=1)
        keys.append(np.minimum(i, j)[ok].astype(np.int64) * n + np.maximum(i, j)[ok])
    for w in np.nonzero(~fin)[0]:
        ok = np.all(bmin[w] <= bmax, axis=1) & np.all(bmin <= bmax[w], axis=1)
        ok[w] = False
        j = np.nonzero(ok)[0]
        keys.append(np.minimum(w, j).astype(np.int64) * n + np.maximum(w, j))
    return np.unique(np.concatenate(keys)) if keys else np.zeros(0, np.int64)


def test_batch_256_specimens_pair_sets(gpu):
    """The batching workload at full size (256 specimens, log-uniform 2,000-50,000 particles,
    3.9M bodies): after the first collider run and after a forced later run, no interaction
    joins two specimens, and every specimen's interactions are exactly its overlapping boxes
    (plus, after the later run, pairs that were real contacts when it ran)."""
    specs = configs.config4(n_specimens=256)
    b = configs.batch(specs)
    bc = b.open(gpu)
    spec = b.arrays["specimen"]
    starts = np.concatenate([[0], np.cumsum([c.n for c in specs])])
    checked = 0
    for run in range(2):
        real_key = np.zeros(0, np.int64)
        if run:
            bc.step(30)
            i1, i2 = bc.get("intr.id1"), bc.get("intr.id2")
            real = (bc.get("intr.hasGeom") != 0) & (bc.get("intr.hasPhys") != 0)
            real_key = np.minimum(i1, i2)[real].astype(np.int64) * b.n + np.maximum(i1, i2)[real]
            bc.set("collider.strideActive", np.array([0], np.int32))
        runs = bc.counters()["collider_runs"]
        bc.step(1)
        assert bc.counters()["collider_runs"] == runs + 1
        i1, i2 = bc.get("intr.id1"), bc.get("intr.id2")
        assert np.array_equal(spec[i1], spec[i2])
        bmin, bmax = bc.get("bound.min").reshape(-1, 3), bc.get("bound.max").reshape(-1, 3)
        lo_id, hi_id = np.minimum(i1, i2).astype(np.int64), np.maximum(i1, i2).astype(np.int64)
        order = np.argsort(spec[i1], kind="stable")
        cut = np.searchsorted(spec[i1][order], np.arange(len(specs) + 1))
        for k in range(len(specs)):
            lo, hi = int(starts[k]), int(starts[k + 1])
            n = hi - lo
            sel = order[cut[k]:cut[k + 1]]
            have = np.sort((lo_id[sel] - lo) * n + (hi_id[sel] - lo))
            assert len(np.unique(have)) == len(have)  # no duplicates
            expect = _box_overlaps(bmin[lo:hi], bmax[lo:hi])
            assert len(np.setdiff1d(expect, have)) == 0, (run, k)
            extra = np.setdiff1d(have, expect)
            extra_global = (extra // n + lo) * b.n + (extra % n + lo)
            assert np.all(np.isin(extra_global, real_key)), (run, k)
            checked += len(expect)
    print("256 specimens: %d overlapping pairs checked over two collider runs" % checked)
