"""Edge cases of the step on the GPU against the oracle.

Scenes of a few bodies (one particle and the six walls) and body counts that fill no launch
block; a step call of zero steps; a scene whose bodies never meet (no interaction with a
particle is ever created); a collider run that inserts a whole packing's interactions at once, so the device's
candidate, hash and interaction buffers grow past their first sizes inside one run.  Bars as in
test_gpu_parity: pair lists (ids, orientation, container order) bit-exact, positions to 1e-12.
"""
import numpy as np
import pytest

from sudodem_amd import configs

pytestmark = pytest.mark.gpu


def _pairs(ctx):
    return np.stack([ctx.get("intr.id1"), ctx.get("intr.id2")], axis=1)


def _same(g, c, what):
    assert np.array_equal(_pairs(g), _pairs(c)), what
    pg, pc = g.get("state.pos"), c.get("state.pos")
    assert np.max(np.abs(pg - pc)) <= 1e-12 * max(1.0, np.max(np.abs(pc))), what
    assert g.counters()["collider_runs"] == c.counters()["collider_runs"], what


def _open_pair(gpu, oracle, cfg):
    g, c = cfg.open(gpu), cfg.open(oracle)
    oracle.set_init_sort(c.ptr, 1)
    return g, c


def _spread(ctx, factor):
    """Walls 1 m away, particle centres scaled by `factor` about the origin."""
    pos = ctx.get("state.pos").reshape(-1, 3)
    for w in range(6):
        pos[w, w // 2] = 1.0 if w % 2 else -1.0
    pos[6:] *= factor
    ctx.set("state.pos", pos)


@pytest.mark.parametrize("n", [1, 2, 7, 65, 257])
def test_small_and_ragged_scenes(gpu, oracle, n):
    g, c = _open_pair(gpu, oracle, configs.config0(n_particles=n))
    for k in range(3):
        g.step(40)
        c.step(40)
        _same(g, c, (n, k))


def test_zero_steps_change_nothing(gpu):
    g = configs.config0(n_particles=50).open(gpu)
    g.step(3)
    before = {k: g.get(k).copy() for k in ("state.pos", "state.vel", "intr.id1", "scene.iter")}
    g.step(0)
    for k, v in before.items():
        assert np.array_equal(g.get(k), v), k


def test_bodies_that_never_meet(gpu, oracle):
    cfg = configs.config0(n_particles=120)
    g, c = _open_pair(gpu, oracle, cfg)
    for ctx in (g, c):
        _spread(ctx, 3.0)  # 8.6 mm between lattice neighbours, particles < 3 mm across
        ctx.step(50)
    _same(g, c, "apart")
    # only the pairs of perpendicular walls, whose plane boxes always overlap (virtual, kept)
    assert np.all(np.maximum(g.get("intr.id1"), g.get("intr.id2")) < 6)
    assert g.counters()["contact_evals"] == 0
    v = g.get("state.vel").reshape(-1, 3)[6:]
    assert np.allclose(v[:, 2], v[0, 2]) and v[0, 2] < 0  # all in the same free fall


def test_whole_packing_inserted_in_one_run(gpu, oracle):
    cfg = configs.config0(n_particles=1000)
    g, c = _open_pair(gpu, oracle, cfg)
    for ctx in (g, c):
        packed = ctx.get("state.pos").copy()
        _spread(ctx, 3.0)
        ctx.step(2)
        ctx.set("state.pos", packed)  # every neighbour pair of the packing at once
        ctx.set("state.vel", np.zeros_like(ctx.get("state.vel")))
    for k in range(3):
        g.step(10)
        c.step(10)
        _same(g, c, k)
    assert len(g.get("intr.id1")) > 2000


def test_observable_argument_errors_device(gpu):
    from test_abi import check_observable_argument_errors
    check_observable_argument_errors(gpu)
