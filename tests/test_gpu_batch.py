"""Batching: several independent specimens advanced in one context give, specimen by
specimen, the result of running each specimen alone (GPU)."""
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
