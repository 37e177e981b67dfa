"""GPU: the HIP frozen-contact rollout and its adjoint against the FP64 CPU forward-mode
restatement (tests/frozen_ref.py) and against central finite differences of the HIP forward, on a
window with loaded contacts (tests/frozen_case.py), with and without local damping and with the
Coulomb cap active; float32 parameters and velocities at the interface."""
import numpy as np
import pytest
import torch

from frozen_case import make_window
from frozen_ref import rollout_jvp
from test_frozen_ref import _fd_agreement

pytestmark = pytest.mark.gpu

STEPS, LOSS_STEPS, DT, CK = 120, 40, 1e-6, 50  # three checkpoint segments, the last one short
G = (0.0, 0.0, -9.81)
CASES = [(0.0, 0.5), (0.3, 0.5), (0.3, 0.05)]  # (local damping, friction coefficient)


@pytest.fixture(scope="module")
def window(gpu, oracle):
    return make_window(oracle)[0]


def _problem(A, damping):
    from sudodem_amd.adjoint import FrozenProblem
    return FrozenProblem(A, DT, damping, G, STEPS, LOSS_STEPS, CK)


def _run(P, params, v0, grad=True, dtype=torch.float64):
    from sudodem_amd.adjoint import rollout
    p = torch.tensor(params, dtype=dtype, device="cuda", requires_grad=grad)
    v = torch.tensor(v0, dtype=dtype, device="cuda", requires_grad=grad)
    loss = rollout(P, p, v)
    if not grad:
        return float(loss), None, None
    loss.backward()
    assert p.grad.dtype == dtype and v.grad.dtype == dtype
    return float(loss.detach()), p.grad.double().cpu().numpy(), v.grad.double().cpu().numpy()


@pytest.mark.parametrize("damping,mu", CASES)
def test_forward_and_adjoint_match_cpu_forward_mode(window, damping, mu):
    A = window
    P = _problem(A, damping)
    params = (1e5, 7e4, mu)
    st = {}
    loss, gp, gv = _run(P, params, A["v0"])
    ref = rollout_jvp(A, params, A["v0"], STEPS, LOSS_STEPS, DT, damping, G, stats=st)[0]
    assert loss > 0 and st["capped"] > 0 and (st["damped"] > 0) == (damping > 0)
    assert loss == pytest.approx(ref, rel=1e-11)
    for k in range(3):
        d = [0.0, 0.0, 0.0]
        d[k] = 1.0
        jvp = rollout_jvp(A, params, A["v0"], STEPS, LOSS_STEPS, DT, damping, G, dparams=d)[1]
        assert jvp != 0
        assert gp[k] == pytest.approx(jvp, rel=1e-8, abs=1e-14 * abs(loss) / params[k])
    rng = np.random.default_rng(3)
    for _ in range(3):
        dv0 = np.zeros_like(A["v0"])
        dv0[A["dynamic"] == 1] = rng.normal(size=(int(A["dynamic"].sum()), 3))
        jvp = rollout_jvp(A, params, A["v0"], STEPS, LOSS_STEPS, DT, damping, G, dv0=dv0)[1]
        assert float(np.sum(gv * dv0)) == pytest.approx(jvp, rel=1e-8)
    assert np.all(gv[A["dynamic"] == 0] == 0)


@pytest.mark.parametrize("damping,mu", CASES)
def test_adjoint_vs_central_differences(window, damping, mu):
    A = window
    P = _problem(A, damping)
    params = (1e5, 7e4, mu)
    loss, gp, gv = _run(P, params, A["v0"])
    for k in range(3):
        def f(h):
            p = list(params)
            p[k] += h
            return _run(P, p, A["v0"], False)[0]
        best, errs = _fd_agreement(f, params[k], gp[k])
        assert best < 1e-4, (k, errs)
    # every velocity component of one body, and a random direction over all bodies
    i = int(np.argmax(np.linalg.norm(gv, axis=1)))
    for c in range(3):
        e = np.zeros_like(A["v0"])
        e[i, c] = 1.0
        best, errs = _fd_agreement(lambda h: _run(P, params, A["v0"] + h * e, False)[0], 1.0, gv[i, c],
                                   rel_steps=(1e-5, 1e-6, 1e-7))
        assert best < 1e-4, (i, c, errs)
    rng = np.random.default_rng(5)
    dv0 = np.zeros_like(A["v0"])
    dv0[A["dynamic"] == 1] = rng.normal(size=(int(A["dynamic"].sum()), 3))
    best, errs = _fd_agreement(lambda h: _run(P, params, A["v0"] + h * dv0, False)[0], 1.0, float(np.sum(gv * dv0)),
                               rel_steps=(1e-5, 1e-6, 1e-7))
    assert best < 1e-4, errs


def test_float32_interface(window):
    """float32 parameters and velocities: the kernels work in double on their float32 values, so
    the loss and gradients are the double ones at the rounded inputs, returned as float32."""
    A = window
    P = _problem(A, 0.3)
    params = (1e5, 7e4, 0.5)
    v32 = A["v0"].astype(np.float32)
    loss32, gp32, gv32 = _run(P, params, v32, dtype=torch.float32)
    loss64, gp64, gv64 = _run(P, params, v32.astype(np.float64))
    assert loss32 == loss64
    assert np.allclose(gp32, gp64, rtol=1e-6, atol=0) and np.allclose(gv32, gv64, rtol=1e-6, atol=1e-6 * np.abs(gv64).max())
