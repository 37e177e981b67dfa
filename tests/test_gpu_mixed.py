"""The mixed FP32/FP64 form against the oracle, per contact.

Each Nelder-Mead distance evaluation is a float difference from a double anchor, taken when
its touching test is certified against its rounding bound and done in double (a new anchor)
otherwise; the simplex stays in double, the vertex comparisons use the float values, and the
depth and points come from a final double evaluation.  Contact classification is then exact
and a comparison within float rounding is the only way to a different solver path.  Bars (one
step from identical states, every contact): the same contacts, overlap within 1e-5 relative
for at least 99.9 % of them and within 1e-8 for 99 %, at least 99 % on the oracle's solver
path, normals within 1e-6 and normal forces within 1e-6 relative for 99.9 % (the curvature
and Hertz terms are double in both forms).
"""
import numpy as np
import pytest

from sudodem_amd import SDM_MIXED, configs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1_bed(gpu):
    return configs.densify(configs.config1(n_particles=3000, box_xy=24e-3), gpu, iterations=60)


def _per_contact(g, c):
    dg, dc = g.get("geom.PenetrationDepth"), c.get("geom.PenetrationDepth")
    both = (dg > 0) & (dc > 0)
    out = {"contacts": int((dc > 0).sum()), "classified_differently": int(((dg > 0) != (dc > 0)).sum()),
           "depth": np.abs(dg[both] - dc[both]) / dc[both]}
    for k, name in (("geom.normal", "normal"), ("phys.normalForce", "normal_force"), ("contact.force", "force")):
        a, b = g.get(k).reshape(-1, 3)[both], c.get(k).reshape(-1, 3)[both]
        out[name] = np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-300)
    out["same_path"] = float(np.mean(g.get("diag.stop")[both] == c.get("diag.stop")[both]))
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
