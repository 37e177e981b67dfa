"""The mixed FP32/FP64 form against the oracle, per contact.

Each Nelder-Mead distance evaluation is a float difference from a double anchor, taken when
its touching test is certified against its rounding bound and done in double (a new anchor)
otherwise; the simplex stays in double, the vertex comparisons use the float values, and the
depth and points come from a final double evaluation.  Contact classification is then exact
and a comparison within float rounding is the only way to a different solver path.  Bars (one
step from identical states, every contact): the same contacts, overlap within 1e-5 relative
for at least 99.9 % of them and within 1e-8 for 99 %, at least 99 % on the oracle's solver
path, normals within 1e-6 and normal forces within 1e-6 relative for 99.9 % (the curvature
and Hertz terms are double in both forms).  The Levenberg-Marquardt solve's mixed form (linear
law, periodic cell and box: double with the device's functions) is checked at the end of this
file.
"""
import numpy as np
import pytest

from sudodem_amd import SDM_MIXED, configs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1_bed(gpu):
    return configs.densify(configs.config1(n_particles=3000, box_xy=24e-3), gpu, iterations=60)


def _per_contact(g, c, by_ids=False):
    """Per-contact differences; by_ids: match the interactions by their body ids (the two
    containers may list them in another order)."""
    og = oc = slice(None)
    if by_ids:
        kg = g.get("intr.id1").astype(np.int64) * (1 << 32) + g.get("intr.id2")
        kc = c.get("intr.id1").astype(np.int64) * (1 << 32) + c.get("intr.id2")
        og, oc = np.argsort(kg), np.argsort(kc)
        assert np.array_equal(kg[og], kc[oc])
    dg, dc = g.get("geom.PenetrationDepth")[og], c.get("geom.PenetrationDepth")[oc]
    both = (dg > 0) & (dc > 0)
    out = {"contacts": int((dc > 0).sum()), "classified_differently": int(((dg > 0) != (dc > 0)).sum()),
           "depth": np.abs(dg[both] - dc[both]) / dc[both]}
    for k, name in (("geom.normal", "normal"), ("phys.normalForce", "normal_force"), ("contact.force", "force")):
        a, b = g.get(k).reshape(-1, 3)[og][both], c.get(k).reshape(-1, 3)[oc][both]
        out[name] = np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-300)
    out["same_path"] = float(np.mean(g.get("diag.stop")[og][both] == c.get("diag.stop")[oc][both]))
    return out


def test_mixed_per_contact_against_oracle(gpu, oracle, c1_bed):
    base = c1_bed.open(gpu)
    base.step(3)
    snap = base.snapshot()
    g = c1_bed.open(gpu, precision=SDM_MIXED)
    g.restore(snap)
    c = c1_bed.open(oracle)
    c.restore(snap)
    g.set_timing(True)
    g.step(1)
    c.step(1)
    assert np.array_equal(g.get("intr.id1"), c.get("intr.id1"))
    kt = g.kernel_timers()
    r = _per_contact(g, c)
    d = r["depth"]
    print("mixed vs oracle: %d contacts, %d classified differently, same path %.4f, %d double evaluations for %d "
          "listed pairs; depth rel err max %.3g, frac <= 1e-5 %.5f; normal max %.3g; normal force max %.3g"
          % (r["contacts"], r["classified_differently"], r["same_path"], kt["nm_double_evals"], kt["nm_listed"],
             d.max(), np.mean(d <= 1e-5), r["normal"].max(), r["normal_force"].max()))
    assert r["contacts"] > 1000
    assert r["classified_differently"] == 0
    assert np.mean(d <= 1e-5) >= 0.999
    assert np.mean(d <= 1e-8) >= 0.99
    assert r["same_path"] >= 0.99
    assert np.mean(r["normal"] <= 1e-6) >= 0.99
    assert np.mean(r["normal_force"] <= 1e-6) >= 0.999
    # the difference form carries the solve: double evaluations are a small share
    assert 0 < kt["nm_listed"] and kt["nm_double_evals"] < 0.5 * kt["nm_listed"]


def _axis_chains():
    """Chains of identically oriented superellipsoids of two alternating shapes along x, y and
    z, each pair overlapping by 2 um: every contact direction starts on a body axis, where the
    difference form admits no anchor (a direction component exactly zero), so the main solver
    launch hands these pairs over and the resume launch solves them."""
    import math
    from sudodem_amd.scene import SuperquadricsMat2
    shapes = [(np.array([1.0e-3, 0.9e-3, 0.8e-3]), (0.8, 1.2)), (np.array([0.85e-3, 1.05e-3, 0.95e-3]), (1.0, 1.1))]
    pos, rxyz, eps = [], [], []
    for axis in range(3):
        x = np.array([6e-3, 6e-3, 6e-3]) + 8e-3 * axis
        for i in range(8):
            r, e = shapes[i % 2]
            if i:
                x = x.copy()
                x[axis] += rxyz[-1][axis] + r[axis] - 2e-6
            pos.append(x)
            rxyz.append(r)
            eps.append(e)
    pos, rxyz, eps = np.array(pos), np.array(rxyz), np.array(eps)
    n = len(pos)
    A = configs._empty_arrays(6 + n)
    lo, hi = np.zeros(3), np.full(3, 60e-3)
    configs._box_walls(A, lo, hi, wall_mat=1)
    quat = np.tile([1.0, 0.0, 0.0, 0.0], (n, 1))
    mat = SuperquadricsMat2(label="sq", G=7e7 / 2.6, nu=0.3, frictionAngle=math.atan(0.5), density=2650.0,
                            Kn=1e5, Ks=7e4)
    wall = SuperquadricsMat2(label="wall", G=7e7 / 2.6, nu=0.3, frictionAngle=math.atan(0.5), density=2650.0,
                             Kn=1e5, Ks=7e4)
    configs._fill_particles(A, 6, pos, rxyz, eps, quat, 0, mat.density)
    return configs.Config("axis-aligned chains", A, [mat, wall], configs.hertz_engines(1e-4), dt=1e-6, steps=1,
                          box=(lo, hi))


def test_mixed_handover_against_oracle(gpu, oracle):
    cfg = _axis_chains()
    base = cfg.open(gpu)
    base.step(1)
    snap = base.snapshot()
    # every solve starts on the line of centres, a body axis of both (unrotated) bodies
    snap["geom.contactAngle"][:] = 0.0
    A = cfg.arrays
    for name, key in (("state.pos", "pos"), ("state.ori", "ori"), ("state.vel", "vel"), ("state.angVel", "angVel"),
                      ("state.angMom", "angMom")):
        snap[name] = A[key].reshape(snap[name].shape).copy()
    for name in ("shape.rot_mat2global", "shape.rot_mat2local"):
        snap[name] = np.tile(np.eye(3).ravel(), cfg.n).reshape(snap[name].shape)
    g = cfg.open(gpu, precision=SDM_MIXED)
    g.restore(snap)
    c = cfg.open(oracle)
    c.restore(snap)
    g.set_timing(True)
    g.step(1)
    c.step(1)
    assert np.array_equal(g.get("intr.id1"), c.get("intr.id1"))
    kt = g.kernel_timers()
    r = _per_contact(g, c)
    print("axis chains: %d contacts, %d handed over, depth rel err max %.3g, same path %.3f"
          % (r["contacts"], kt["nm_handed_over"], r["depth"].max(), r["same_path"]))
    assert r["contacts"] >= 20 and r["classified_differently"] == 0
    assert kt["nm_handed_over"] >= 7
    assert r["depth"].max() <= 1e-8
    # the chain along x has every direction component exactly zero where it is not one (its
    # CRotationMat frame is the identity): each of its solves is the oracle's.  Along y and z
    # the frame carries a 2e-16 component, so the solves re-anchor at every point, and in
    # these mirror-symmetric pairs vertex values tie to within that rounding: a tie the
    # other way changes the last evaluated point, hence the normal, not the best vertex.
    assert np.mean(r["normal"] <= 1e-6) >= 0.3 and np.mean(r["normal_force"] <= 1e-6) >= 0.3


def _linear_bed(gpu):
    import math
    from sudodem_amd.scene import SuperquadricsMat
    cfg = configs.config1(n_particles=1000, box_xy=20e-3)
    cfg.engines = configs.linear_engines(cfg.engines[1].verletDist)
    cfg.materials = [SuperquadricsMat(label="mat1", Kn=1e5, Ks=7e4, frictionAngle=math.atan(0.5), density=2650.0),
                     SuperquadricsMat(label="wallmat", Kn=1e5, Ks=7e4, frictionAngle=0.0, density=2650.0)]
    return configs.densify(cfg, gpu, iterations=40)


@pytest.mark.parametrize("case", ["periodic", "periodic_shaken", "linear_bed"])
def test_mixed_levenberg_marquardt_against_oracle(gpu, oracle, case):
    """The mixed form of the Levenberg-Marquardt solve (Superquadrics_Ig2.cpp:1056-1354): the
    certified float screen of the first evaluation, then the solve in double with the device's
    own functions (not the correctly rounded ones of the FP64 form).  Bars from identical
    states, one step, every contact: no contact classified differently, the oracle's solver
    path for 99 %, depth within 1e-5 relative for 99 % (measured 99.8 % periodic: the rest are
    solves that a last-bit difference of a device function sends to another end point; 100 %
    bed)."""
    if case == "linear_bed":
        cfg, warm, vel = _linear_bed(gpu), 2, None
    else:
        cfg = configs.densify(configs.config2(n_particles=2000), gpu, iterations=80, growth=0.01, relax=20)
        warm = 3
        vel = np.random.default_rng(1).normal(scale=1.0, size=(cfg.n, 3)) if case == "periodic_shaken" else None
    c = cfg.open(oracle)  # the state (rings included) comes from the oracle
    if vel is not None:
        c.set("state.vel", vel)
    c.step(warm)
    g = cfg.open(gpu, precision=SDM_MIXED)
    g.restore(c.snapshot())
    g.set_timing(True)
    g.step(1)
    c.step(1)
    assert np.array_equal(g.get("intr.id1"), c.get("intr.id1")) and np.array_equal(g.get("intr.id2"), c.get("intr.id2"))
    ndbl, nev = g.kernel_timers()["nm_double_evals"], g.timers()["solver_evals"]
    r = _per_contact(g, c)
    d = r["depth"]
    print("%s: %d contacts, %d classified differently, same path %.4f, depth rel err p99 %.3g max %.3g, "
          "frac <= 1e-5 %.5f; %d of %d evaluations in double"
          % (case, r["contacts"], r["classified_differently"], r["same_path"], np.percentile(d, 99), d.max(),
             np.mean(d <= 1e-5), ndbl, nev))
    assert r["contacts"] > 500
    assert r["classified_differently"] == 0
    assert r["same_path"] >= 0.99
    assert np.mean(d <= 1e-5) >= 0.99
    assert np.mean(r["normal_force"] <= 1e-5) >= 0.99
    assert ndbl >= nev > 0  # every evaluation in double (restarts included in ndbl)
