"""The hand-computed adaptive-sweep and swap-with-last vectors of
tests/test_oracle_hand_vectors.py, on the device."""
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
