"""Wall stress servo on the GPU (libsdm_hip.so) against the oracle: consolidation,
hydroConsolidation, drained and constant-volume shear, the echo.

Bars: the servo's state and the wall positions follow the oracle's through gain updates,
stability checks and completion (FP64 step; wall forces are ordered sums, stiffness sums are in
interaction order, so stiffness is bit-exact and the rest agrees to 1e-9 relative); the check's
tensor and unbalanced force are block-tree sums on the device (1e-12 relative); completion
happens at the same iteration, quiets the system and ends the step call; steps replayed from
the captured step graph give the state of launched steps bit for bit.
"""
import numpy as np
import pytest

import test_servo as TS

pytestmark = pytest.mark.gpu

SERVO = dict(goal=(2e4, 2e4, 2e4), max_vel=0.05, gain_alpha=0.5, iterate_num_max=5, solve_num_max=20,
             f_threshold=0.01, fmin_threshold=0.01, unbf_tol=0.01)


def _pair(gpu, oracle, **servo):
    # walls resting on the outermost particles, which are far apart: wall contacts only, so the
    # two steps stay within rounding of each other
    out = []
    for lib in (gpu, oracle):
        cfg, ctx = TS.gentle_scene(lib)
        ctx.step(3)
        ctx.set_servo(**servo)
        out.append(ctx)
    return out


def _agree(g, c, what):
    sg, sc = g.servo_state(), c.servo_state()
    assert np.array_equal(sg["stiffness"], sc["stiffness"]), what
    for k in ("wall_stress", "stress", "gain", "size"):
        a, b = np.asarray(sg[k], float), np.asarray(sc[k], float)
        fin = np.isfinite(b)
        assert np.array_equal(np.isfinite(a), fin), (what, k)
        assert np.allclose(a[fin], b[fin], rtol=1e-9, atol=1e-9 * np.abs(b[fin]).max(initial=1.0)), (what, k, a, b)
    for k in ("iterate_num", "solve_num", "completed_iter"):
        assert sg[k] == sc[k], (what, k)
    for k in ("mean_stress", "deviator_stress", "unbalanced_force"):
        assert np.isclose(sg[k], sc[k], rtol=1e-9, atol=1e-12), (what, k, sg[k], sc[k])
    wg = g.get("state.pos").reshape(-1, 3)[:6]
    wc = c.get("state.pos").reshape(-1, 3)[:6]
    assert np.allclose(wg, wc, rtol=0, atol=1e-12), what


def test_servo_follows_oracle(gpu, oracle):
    g, c = _pair(gpu, oracle, **SERVO)
    ws_max, mean_max = np.zeros(6), 0.0
    for k in range(6):
        g.step(10)
        c.step(10)
        _agree(g, c, "after %d steps" % (10 * (k + 1)))
        st = g.servo_state()
        ws_max = np.maximum(ws_max, st["wall_stress"])
        mean_max = max(mean_max, st["mean_stress"])
    assert np.all(st["stiffness"] > 0) and np.all(ws_max > 0)  # every wall pushed
    assert mean_max > 0  # checks ran on contacts


def test_completion_same_iteration(gpu, oracle):
    g, c = _pair(gpu, oracle, **{**SERVO, "f_threshold": 1e9, "fmin_threshold": 1e9, "unbf_tol": 1e9})
    for ctx in (g, c):
        ctx.step(100)  # the first check (20th servo step) completes
    assert int(g.get("scene.iter")[0]) == int(c.get("scene.iter")[0]) == 3 + 20
    _agree(g, c, "completion")
    assert g.servo_state()["completed_iter"] == 3 + 19
    assert np.all(g.get("state.vel") == 0) and np.all(g.get("state.angVel") == 0)


def test_servo_in_step_graph(gpu):
    # fixed walls (the servo still updates gains and runs its checks): the collider settles
    # and steps replay the captured graph with the servo kernels in it
    servo = {**SERVO, "activated": (0, 0, 0, 0), "z_servo": 0, "iterate_num_max": 3, "solve_num_max": 7}
    res = []
    for graph in (True, False):
        cfg, ctx = TS.gentle_scene(gpu, overlap=-1e-4)  # walls 0.1 mm clear of the particles
        ctx.set("newton.gravity", np.zeros(3))
        ctx.set_graph(graph)
        ctx.step(3)
        ctx.set_servo(**servo)
        ctx.step(40)
        res.append((ctx.servo_state(), ctx.get("state.pos"), ctx.counters()["graph_steps"]))
    (sa, pa, ga), (sb, pb, gb) = res
    assert ga >= 10 and gb == 0
    assert np.array_equal(pa, pb)
    for k in sa:
        assert np.array_equal(np.asarray(sa[k]), np.asarray(sb[k])), k


def _agree_more(g, c, what):
    _agree(g, c, what)
    sg, sc = g.servo_state(), c.servo_state()
    assert sg["ramp_num"] == sc["ramp_num"], what
    assert np.array_equal(sg["size0"], sc["size0"]), what


def test_shear_after_consolidation_follows_oracle(gpu, oracle):
    # consolidation, then the same engine switched to drained shear (ramp of example2_shear.py)
    g, c = _pair(gpu, oracle, **SERVO, echo_interval=7)
    for ctx in (g, c):
        ctx.step(12)
    _agree_more(g, c, "consolidated")
    dt = float(c.get("scene.dt")[0])
    shear = {**SERVO, "goal": (2e4, 2e4, 2e-4 / dt), "z_servo": 0, "keep_state": 1, "ramp_interval": 10,
             "ramp_chunks": 2, "echo_interval": 7, "iterate_num_max": 4}
    for ctx in (g, c):
        ctx.set_servo(**shear)
    for k in range(4):
        g.step(6)
        c.step(6)
        _agree_more(g, c, "shear, %d steps" % (6 * (k + 1)))
    assert g.servo_state()["ramp_num"] == 10


def test_constant_volume_shear_ends_at_same_iteration(gpu, oracle):
    g, c = _pair(gpu, oracle, **SERVO)
    dt = float(c.get("scene.dt")[0])
    shear = {**SERVO, "goal": (2e4, 2e4, 1e-3 / dt), "z_servo": 0, "shear_mode": 2, "ramp_interval": 4,
             "ramp_chunks": 2, "echo_interval": 3, "target_strain": 0.01}
    for ctx in (g, c):
        ctx.set_servo(**shear)
        ctx.step(100)
    assert int(g.get("scene.iter")[0]) == int(c.get("scene.iter")[0]) < 3 + 100
    _agree_more(g, c, "end of shear")
    assert g.servo_state()["completed_iter"] >= 0


def test_hydro_follows_oracle(gpu, oracle):
    hydro = {**SERVO, "hydro_strain": 1, "hydro_strain_rate": 5.0, "solve_num_max": 6, "iterate_num_max": 4,
             "echo_interval": 5, "f_threshold": 1e9, "fmin_threshold": 1e9, "unbf_tol": 1e9}
    g, c = _pair(gpu, oracle, **hydro)
    for k in range(3):
        g.step(4)
        c.step(4)
        _agree_more(g, c, "hydro, %d steps" % (4 * (k + 1)))
    for ctx in (g, c):
        ctx.step(100)
    assert int(g.get("scene.iter")[0]) == int(c.get("scene.iter")[0]) < 3 + 112
    _agree_more(g, c, "hydro completion")
    assert np.any(g.get("state.vel") != 0)
