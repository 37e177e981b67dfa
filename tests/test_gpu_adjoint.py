"""GPU: the HIP frozen-contact rollout and its adjoint against the FP64 CPU forward-mode
restatement (tests/frozen_ref.py) and against central finite differences of the HIP forward."""
import numpy as np
import pytest
import torch

from frozen_case import make_arrays
from frozen_ref import rollout_jvp

pytestmark = pytest.mark.gpu

STEPS, LOSS_STEPS, DT, CK = 120, 40, 1e-6, 50  # three checkpoint segments, the last one short
PARAMS = (1e5, 7e4, 0.5)
G = (0.0, 0.0, -9.81)


@pytest.fixture(scope="module")
def case(gpu, oracle):
    from sudodem_amd.adjoint import FrozenProblem
    A = make_arrays(oracle)
    P = FrozenProblem(A, DT, 0.0, G, STEPS, LOSS_STEPS, CK)
    return A, P


def _run(P, params, v0, grad=True):
    from sudodem_amd.adjoint import rollout
    p = torch.tensor(params, dtype=torch.float64, device="cuda", requires_grad=grad)
    v = torch.tensor(v0, dtype=torch.float64, device="cuda", requires_grad=grad)
    loss = rollout(P, p, v)
    if not grad:
        return float(loss), None, None
    loss.backward()
    return float(loss.detach()), p.grad.cpu().numpy(), v.grad.cpu().numpy()


def test_forward_and_adjoint_match_cpu_forward_mode(case):
    A, P = case
    loss, gp, gv = _run(P, PARAMS, A["v0"])
    ref = rollout_jvp(A, PARAMS, A["v0"], STEPS, LOSS_STEPS, DT, 0.0, G)[0]
    assert loss == pytest.approx(ref, rel=1e-11)
    for k in range(3):
        d = [0.0, 0.0, 0.0]
        d[k] = 1.0
        jvp = rollout_jvp(A, PARAMS, A["v0"], STEPS, LOSS_STEPS, DT, 0.0, G, dparams=d)[1]
        assert gp[k] == pytest.approx(jvp, rel=1e-8, abs=1e-14 * abs(loss) / PARAMS[k])
    rng = np.random.default_rng(3)
    for _ in range(3):
        dv0 = np.zeros_like(A["v0"])
        dv0[A["dynamic"] == 1] = rng.normal(size=(int(A["dynamic"].sum()), 3))
        jvp = rollout_jvp(A, PARAMS, A["v0"], STEPS, LOSS_STEPS, DT, 0.0, G, dv0=dv0)[1]
        assert float(np.sum(gv * dv0)) == pytest.approx(jvp, rel=1e-8)
    assert np.all(gv[A["dynamic"] == 0] == 0)


def test_adjoint_vs_central_differences(case):
    A, P = case
    loss, gp, gv = _run(P, PARAMS, A["v0"])
    for k in range(3):
        h = 1e-6 * PARAMS[k]
        p1, p0 = list(PARAMS), list(PARAMS)
        p1[k] += h
        p0[k] -= h
        fd = (_run(P, p1, A["v0"], False)[0] - _run(P, p0, A["v0"], False)[0]) / (2 * h)
        assert gp[k] == pytest.approx(fd, rel=1e-4, abs=1e-10 * abs(loss) / PARAMS[k])
    i = int(np.argmax(np.abs(gv[:, 2])))
    h = 1e-7
    v1, v0 = A["v0"].copy(), A["v0"].copy()
    v1[i, 2] += h
    v0[i, 2] -= h
    fd = (_run(P, PARAMS, v1, False)[0] - _run(P, PARAMS, v0, False)[0]) / (2 * h)
    assert gv[i, 2] == pytest.approx(fd, rel=1e-4)
