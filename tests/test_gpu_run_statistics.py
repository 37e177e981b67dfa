"""A multi-step run on the GPU against the oracle from one state: a compressed 3,000-particle
Hertz-Mindlin bed (walls, gravity) relaxes its overlaps over 200 steps (the mean coordination
number falls from ~2 to ~0.9).  Single contacts of the mixed form have left the FP64 path by
then, but the assembly's state agrees within 1 %: mean coordination number, mean stress of
the Love-Weber tensor and height of the centre of mass, for the FP64 form and the mixed form
(measured: FP64 CN identical, stress 1e-5, height 1e-11; mixed CN 0.3 %, stress 0.06 %)."""
import numpy as np
import pytest

from sudodem_amd import SDM_MIXED, configs

pytestmark = pytest.mark.gpu


def _state(ctx):
    pos = ctx.get("state.pos").reshape(-1, 3)[6:]
    return {"cn": ctx.mean_coordination(), "p": -np.trace(ctx.stress()) / 3.0, "zc": pos[:, 2].mean()}


def test_run_statistics_match_oracle(gpu, oracle):
    # a 3,000-particle column (24x24 mm footprint) densified into a contact-carrying bed
    cfg = configs.densify(configs.config1(n_particles=3000, box_xy=24e-3), gpu, iterations=60)
    c = cfg.open(oracle)
    oracle.set_init_sort(c.ptr, 1)
    c.step(3)
    snap = c.snapshot()
    runs = {}
    for name, prec in (("fp64", None), ("mixed", SDM_MIXED)):
        g = cfg.open(gpu) if prec is None else cfg.open(gpu, precision=prec)
        g.restore(snap)
        runs[name] = g
    for ctx in [c, *runs.values()]:
        ctx.step(200)
    ref = _state(c)
    assert ref["cn"] > 0.5 and ref["p"] > 0
    for name, g in runs.items():
        st = _state(g)
        print(name, {k: (float(st[k]), float(ref[k])) for k in st})
        for k in st:
            assert abs(st[k] - ref[k]) <= 0.01 * abs(ref[k]), (name, k, st[k], ref[k])
