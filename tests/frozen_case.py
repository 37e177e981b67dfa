"""A small frozen-contact window shared by the adjoint tests: 150 superellipsoids settled in a
box by the wall servo (configs.config3: 20 kPa on every wall, 6,000 steps of the linear law with
gravity and local damping 0.3), then the top wall driven down at 0.05 m/s.  The window carries
its load: the top wall's normal force grows over the window, some contacts sit on the Coulomb
cap and the Cundall factors switch."""
import numpy as np

from sudodem_amd import configs
from sudodem_amd.adjoint import FrozenProblem

TOP_WALL = configs.TOP_WALL  # z max wall (sense -1)
WALL_VEL = (0.0, 0.0, -0.05)


def _by_ids(ctx, ids_a, ids_b, name):
    """Rows of an interaction attribute for the given (id1, id2) pairs."""
    i1, i2 = ctx.get("intr.id1"), ctx.get("intr.id2")
    key = {(int(a), int(b)): k for k, (a, b) in enumerate(zip(i1, i2))}
    rows = np.array([key[(int(a), int(b))] for a, b in zip(ids_a, ids_b)])
    return ctx.get(name).reshape(len(i1), -1)[rows]


def make_window(lib, n=150, box_xy=9e-3, settle_steps=6000):
    """The frozen window's arrays and, for its contacts, the engine's normal and shear forces of
    the last step before the window (SuperquadricsLaw::go) and of one further engine step."""
    ctx, cfg = configs.config3(lib, n_particles=n, box_xy=box_xy, settle_steps=settle_steps)
    try:
        A = FrozenProblem.arrays_from_context(ctx, loss_body=TOP_WALL, wall_velocity={TOP_WALL: WALL_VEL})
        eng = {"fn": _by_ids(ctx, A["id_a"], A["id_b"], "phys.normalForce"),
               "fs": _by_ids(ctx, A["id_a"], A["id_b"], "phys.shearForce"),
               "kn": _by_ids(ctx, A["id_a"], A["id_b"], "phys.kn")[:, 0],
               "ks": _by_ids(ctx, A["id_a"], A["id_b"], "phys.ks")[:, 0], "dt": cfg.dt}
        ctx.step(1)
        eng["fn_next"] = _by_ids(ctx, A["id_a"], A["id_b"], "phys.normalForce")
        eng["fs_next"] = _by_ids(ctx, A["id_a"], A["id_b"], "phys.shearForce")
    finally:
        ctx.close()
    return A, eng


def make_arrays(lib):
    return make_window(lib)[0]
