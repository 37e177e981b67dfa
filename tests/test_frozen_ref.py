"""CPU: the forward-mode restatement of the frozen-contact rollout (tests/frozen_ref.py)
against central finite differences of itself, on a small window with real contacts."""
import numpy as np
import pytest

from frozen_case import make_arrays
from frozen_ref import rollout_jvp

STEPS, LOSS_STEPS, DT = 30, 10, 1e-6
PARAMS = (1e5, 7e4, 0.5)


@pytest.fixture(scope="module")
def arrays(oracle):
    return make_arrays(oracle)


def _loss(A, params, v0):
    return rollout_jvp(A, params, v0, STEPS, LOSS_STEPS, DT, 0.0, (0, 0, -9.81))[0]


def test_window_has_loss_contacts(arrays):
    assert len(arrays["depth0"]) > 50
    assert arrays["in_loss"].sum() >= 3


@pytest.mark.parametrize("k", [0, 1, 2])
def test_param_tangent_vs_central_difference(arrays, k):
    A = arrays
    d = [0.0, 0.0, 0.0]
    d[k] = 1.0
    loss, dl = rollout_jvp(A, PARAMS, A["v0"], STEPS, LOSS_STEPS, DT, 0.0, (0, 0, -9.81), dparams=d)
    h = 1e-6 * PARAMS[k]
    p1, p0 = list(PARAMS), list(PARAMS)
    p1[k] += h
    p0[k] -= h
    fd = (_loss(A, p1, A["v0"]) - _loss(A, p0, A["v0"])) / (2 * h)
    assert loss > 0
    assert dl == pytest.approx(fd, rel=1e-5, abs=1e-12 * abs(loss) / PARAMS[k])


def test_velocity_tangent_vs_central_difference(arrays):
    A = arrays
    rng = np.random.default_rng(7)
    dv0 = np.zeros_like(A["v0"])
    dv0[A["dynamic"] == 1] = rng.normal(size=(int(A["dynamic"].sum()), 3))
    loss, dl = rollout_jvp(A, PARAMS, A["v0"], STEPS, LOSS_STEPS, DT, 0.0, (0, 0, -9.81), dv0=dv0)
    h = 1e-7
    fd = (_loss(A, PARAMS, A["v0"] + h * dv0) - _loss(A, PARAMS, A["v0"] - h * dv0)) / (2 * h)
    assert dl == pytest.approx(fd, rel=1e-5)
