"""The captured step graph (sdm_set_graph) replays exactly what the launched step computes.

Each case runs the same configuration twice on the GPU, once with steady-state steps replayed
from the captured graph and once with every kernel launched, and requires the complete state
(bodies, interactions, geometry, physics) to be bitwise identical, with most steps replayed.
"""
import numpy as np
import pytest

from sudodem_amd import SDM_FP64, SDM_MIXED, configs

pytestmark = pytest.mark.gpu

FIELDS = ("state.pos", "state.ori", "state.vel", "state.angVel", "intr.id1", "intr.id2", "intr.iterMadeReal",
          "geom.PenetrationDepth", "geom.normal", "geom.contactAngle", "phys.shearForce", "phys.normalForce",
          "phys.kn", "contact.force", "diag.nEval")


def _run_twice(gpu, cfg, steps, precision=SDM_FP64):
    out = []
    for graph in (True, False):
        ctx = cfg.open(gpu, precision=precision)
        ctx.set_graph(graph)
        ctx.step(steps)
        out.append(({k: ctx.get(k) for k in FIELDS}, ctx.counters()))
        ctx.close()
    (a, ca), (b, cb) = out
    print("%s: %d steps, %d replayed, %d collider runs" % (cfg.name[:40], steps, ca["graph_steps"],
                                                           ca["collider_runs"]), flush=True)
    assert cb["graph_steps"] == 0
    assert ca["graph_steps"] >= 20
    assert ca["collider_runs"] == cb["collider_runs"]
    for k in FIELDS:
        assert np.array_equal(a[k], b[k], equal_nan=True), k


def test_graph_linear_law_walls(gpu):
    _run_twice(gpu, configs.config0(n_particles=300), 150)


@pytest.mark.parametrize("precision", [SDM_FP64, SDM_MIXED])
def test_graph_hertz(gpu, precision):
    cfg = configs.densify(configs.config1(n_particles=3000, box_xy=24e-3), gpu, iterations=40)
    _run_twice(gpu, cfg, 120, precision)


def test_graph_periodic(gpu):
    cfg = configs.densify(configs.config2(n_particles=1500), gpu, iterations=40, growth=0.01, relax=10)
    _run_twice(gpu, cfg, 120)


def test_graph_batch(gpu):
    rng = np.random.default_rng(5)
    specs = [configs.box_specimen(n, rng, kn, mu) for n, kn, mu in ((300, 1e5, 0.5), (700, 6e4, 0.3), (200, 2e5, 0.7))]
    _run_twice(gpu, configs.batch(specs), 150)


def test_tuning_changes_no_result(gpu):
    """Launch shapes (sdm_set_tuning, chosen by sudodem_amd/autotune.py) only reschedule work."""
    cfg = configs.densify(configs.config1(n_particles=2000, box_xy=20e-3), gpu, iterations=30)
    ref = None
    for tune in ({}, {"solver_blocks_per_cu": 1, "sort_by_cost": 0}, {"solver_blocks_per_cu": 4, "graph": 0,
                                                                      "misc_blocks_per_cu": 1}):
        ctx = cfg.open(gpu, precision=SDM_MIXED)
        ctx.set_tuning(**tune)
        ctx.step(60)
        out = {k: ctx.get(k) for k in FIELDS}
        ctx.close()
        if ref is None:
            ref = out
        else:
            for k in FIELDS:
                assert np.array_equal(out[k], ref[k], equal_nan=True), (tune, k)
