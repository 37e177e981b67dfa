"""CPU: the forward-mode restatement of the frozen-contact rollout (tests/frozen_ref.py) against
central finite differences of itself, and its first step against the engine's linear law, on
a window with loaded contacts (tests/frozen_case.py: a servo-settled bed under a moving wall)."""
import numpy as np
import pytest

from frozen_case import make_window
from frozen_ref import rollout_jvp

STEPS, LOSS_STEPS, DT = 30, 10, 1e-6
G = (0, 0, -9.81)
CASES = [(0.0, 0.5), (0.3, 0.5), (0.3, 0.05)]  # (local damping, friction coefficient)


@pytest.fixture(scope="module")
def window(oracle):
    return make_window(oracle)


def _loss(A, params, v0, damping):
    return rollout_jvp(A, params, v0, STEPS, LOSS_STEPS, DT, damping, G)[0]


def test_window_carries_load(window):
    A, _ = window
    assert len(A["depth0"]) > 100 and A["in_loss"].sum() >= 1
    for damping, mu in CASES:
        st = {}
        loss = rollout_jvp(A, (1e5, 7e4, mu), A["v0"], 120, 40, DT, damping, G, stats=st)[0]
        assert loss > 0 and st["capped"] > 0  # the Coulomb cap is active in every case
        assert (st["damped"] > 0) == (damping > 0)  # the Cundall factors act with damping only


def test_first_step_forces_equal_engine_law(window):
    """The window's first step against SuperquadricsLaw::go (pkg/dem/Superquadrics.cpp:1181-1363):
    its normal forces are the engine's of the last step before the window (kn depth n, same
    depth and normal: equal to rounding); its shear forces on contacts off the Coulomb cap are
    the engine's next step to first order (the engine also rotates the shear force with the
    normal and takes the relative velocity at the new contact point)."""
    A, eng = window
    assert np.all(eng["kn"] == 1e5) and np.all(eng["ks"] == 7e4)
    st = {}
    rollout_jvp(A, (1e5, 7e4, 0.5), A["v0"], 1, 1, DT, 0.3, G, stats=st)
    fn = np.linalg.norm(eng["fn"], axis=1)
    assert np.all(np.linalg.norm(st["fn1"] - eng["fn"], axis=1) <= 1e-14 * fn)
    inc = np.linalg.norm(eng["fs_next"] - eng["fs"], axis=1)
    fsn, fnn = np.linalg.norm(eng["fs_next"], axis=1), np.linalg.norm(eng["fn_next"], axis=1)
    fs1, fn1 = np.linalg.norm(st["fs1"], axis=1), np.linalg.norm(st["fn1"], axis=1)
    off_cap = (fsn < 0.99 * 0.5 * fnn) & (fs1 < 0.99 * 0.5 * fn1)
    assert off_cap.sum() > 50
    diff = np.linalg.norm(st["fs1"] - eng["fs_next"], axis=1)
    assert np.all(diff[off_cap] <= 0.2 * inc[off_cap] + 1e-12)
    assert np.median(diff[off_cap] / inc[off_cap]) < 0.01


def _fd_agreement(f, x0, tangent, rel_steps=(1e-4, 1e-5, 1e-6)):
    """Smallest relative difference between the tangent and central differences of f(h) (the
    loss at a step h) over the step sizes r x0: the loss is piecewise smooth -- a contact leaving the window
    drops its shear force, a body's Cundall factor switches sign -- so a step that straddles
    such an event is off by the jump over the step, however small the step; the tangent holds
    the branches at their values and must agree with the differences of a step that straddles
    none."""
    errs = []
    for r in rel_steps:
        h = r * x0
        fd = (f(h) - f(-h)) / (2 * h)
        errs.append(abs(fd - tangent) / abs(tangent))
    return min(errs), errs


@pytest.mark.parametrize("damping,mu", CASES)
@pytest.mark.parametrize("k", [0, 1, 2])
def test_param_tangent_vs_central_difference(window, damping, mu, k):
    A, _ = window
    params = (1e5, 7e4, mu)
    d = [0.0, 0.0, 0.0]
    d[k] = 1.0
    loss, dl = rollout_jvp(A, params, A["v0"], STEPS, LOSS_STEPS, DT, damping, G, dparams=d)
    assert loss > 0 and dl != 0

    def f(h):
        p = list(params)
        p[k] += h
        return _loss(A, p, A["v0"], damping)
    best, errs = _fd_agreement(f, params[k], dl)
    assert best < 1e-5, errs


@pytest.mark.parametrize("damping,mu", CASES)
def test_velocity_tangent_vs_central_difference(window, damping, mu):
    A, _ = window
    params = (1e5, 7e4, mu)
    rng = np.random.default_rng(7)
    dv0 = np.zeros_like(A["v0"])
    dv0[A["dynamic"] == 1] = rng.normal(size=(int(A["dynamic"].sum()), 3))
    loss, dl = rollout_jvp(A, params, A["v0"], STEPS, LOSS_STEPS, DT, damping, G, dv0=dv0)
    best, errs = _fd_agreement(lambda h: _loss(A, params, A["v0"] + h * dv0, damping), 1.0, dl,
                               rel_steps=(1e-5, 1e-6, 1e-7))
    assert best < 1e-5, errs
