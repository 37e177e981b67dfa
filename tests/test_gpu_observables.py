"""Observables on the GPU against the oracle on identical states: the mean coordination number
exactly, the Love-Weber stress to 1e-12 relative (block-tree against serial summation), wall
forces after one step to the per-step parity bar."""
import numpy as np
import pytest

from sudodem_amd import configs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("which", ["walls", "periodic"])
def test_observables_match_oracle(gpu, oracle, which):
    if which == "walls":
        cfg = configs.densify(configs.config1(n_particles=2000, box_xy=20e-3), gpu, iterations=30)
    else:
        cfg = configs.densify(configs.config2(n_particles=1500), gpu, iterations=40, growth=0.01, relax=10)
    c = cfg.open(oracle)
    oracle.set_init_sort(c.ptr, 1)
    c.step(5)
    g = cfg.open(gpu)
    g.restore(c.snapshot())
    assert g.mean_coordination() == c.mean_coordination()
    sc, sg = c.stress(), g.stress()
    assert np.abs(sc).max() > 0
    assert np.allclose(sg, sc, rtol=0, atol=1e-12 * np.abs(sc).max())
    c.step(1)
    g.step(1)
    ids = list(range(6)) if which == "walls" else [0, 7, 100]
    fc, fg = c.body_forces(ids), g.body_forces(ids)
    assert np.allclose(fg, fc, rtol=0, atol=1e-9 * max(np.abs(fc).max(), 1e-30))
