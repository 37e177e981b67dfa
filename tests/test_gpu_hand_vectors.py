"""The hand-computed adaptive-sweep and swap-with-last vectors of
tests/test_oracle_hand_vectors.py and the closed-form solver outputs of
tests/test_oracle_solver_quirks.py, on the device."""
import pytest

import test_oracle_hand_vectors as H

pytestmark = pytest.mark.gpu


def test_adaptive_sweep_constant_velocity_device(gpu):
    H.test_adaptive_sweep_constant_velocity(gpu)


def test_adaptive_sweep_decay_and_cap_device(gpu):
    H.test_adaptive_sweep_decay_and_cap(gpu)


@pytest.mark.parametrize("vel,first,after", H.ERASE_CASES)
def test_swap_with_last_erase_device(gpu, vel, first, after):
    H.test_swap_with_last_erase(gpu, vel, first, after)


def test_lm_output_quirks_closed_form_device(gpu):
    import test_oracle_solver_quirks as Q
    Q.check_lm_outputs(Q.ellipsoid_bed(gpu, "linear"))


def test_nm_output_quirks_closed_form_device(gpu):
    import test_oracle_solver_quirks as Q
    Q.check_nm_outputs(Q.ellipsoid_bed(gpu, "hertz"))
