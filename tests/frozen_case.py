"""A small frozen-contact window shared by the adjoint tests: 150 superellipsoids densified
into a bed (oracle engine), the top wall lowered onto it, one step to form the contacts."""
import numpy as np

from sudodem_amd import configs
from sudodem_amd.adjoint import FrozenProblem
from sudodem_amd.scene import rot_matrices

TOP_WALL = 5  # z max wall (sense -1)
WALL_VEL = (0.0, 0.0, -0.05)


def make_arrays(lib, n=150, box_xy=9e-3):
    cfg = configs.densify(configs.config1(n_particles=n, box_xy=box_xy), lib, iterations=40)
    A = cfg.arrays
    part = A["kind"] == 0
    ez = np.zeros((part.sum(), 3))
    ez[:, 2] = 1.0
    h = configs.support_h(A["rxyz"][part], A["eps"][part], rot_matrices(A["ori"][part]), ez)
    A["pos"][TOP_WALL, 2] = float(np.sort(A["pos"][part, 2] + h)[-4]) - 1e-6  # the 4 tallest touch it
    ctx = cfg.open(lib)
    ctx.set("scene.dt", np.array([1e-7]))
    ctx.step(1)
    arrays = FrozenProblem.arrays_from_context(ctx, loss_body=TOP_WALL, wall_velocity={TOP_WALL: WALL_VEL})
    ctx.close()
    return arrays
