"""Seeded synthetic superellipsoid packings for the benchmark and the parity tests.

Every generator draws from numpy.random.Generator(PCG64(seed)) with seed 20260000 + the
configuration index, creates the six walls first (ids 0-5, which the reference's Ip2/Law2
functors require, pkg/dem/Superquadrics.cpp:1100,1134,1394) and then the particles the way
NewSuperquadrics_rot2 does (py/_superquadrics_utils.cpp:354-379): mass = density * volume,
inertia = density * principal inertia, orientation = normalised random quaternion.

Initial overlaps are controlled with the exact support function of each superellipsoid along
the line of centres (Normal2Phi + getSurface): for convex bodies the overlap along any one
direction bounds the penetration depth from above, so bounding the centre-line overlap bounds
the initial contact overlaps.
"""
from __future__ import annotations

import copy
import math

import numpy as np

from . import _native as nat
from .scene import (BODY_ARRAYS, Bo1_Superquadrics_Aabb, Bo1_Wall_Aabb, ForceResetter,
                    Ig2_Superquadrics_Superquadrics_SuperquadricsGeom, Ig2_Superquadrics_Superquadrics_SuperquadricsGeom2,
                    Ig2_Wall_Superquadrics_SuperquadricsGeom, Ig2_Wall_Superquadrics_SuperquadricsGeom2,
                    InsertionSortCollider, InteractionLoop, Ip2_SuperquadricsMat2_SuperquadricsMat2_HertzMindlinPhys,
                    Ip2_SuperquadricsMat_SuperquadricsMat_SuperquadricsPhys, NewtonIntegrator, SuperquadricsLaw,
                    SuperquadricsLaw2, SuperquadricsMat, SuperquadricsMat2, engines_to_params, rot_matrices)

SEED_BASE = 20260000
DOF_ALL = 63


def rng_for(config_index: int, extra: int = 0):
    return np.random.Generator(np.random.PCG64(SEED_BASE + config_index + 1000 * extra))


def random_quaternions(rng, n):
    q = rng.normal(size=(n, 4))
    return q / np.linalg.norm(q, axis=1, keepdims=True)


def superquadric_props(rxyz, eps):
    """Vectorised front end of sdm_superquadric_props (Superquadrics::Initial)."""
    return nat.load_hip().props_n(rxyz, eps)


def support_h(rxyz, eps, R, d):
    """Support function h(d) = max_{p in body} p.d for unit directions d (rows), bodies
    given by semi-axes, exponents and rotation matrices to the global frame (rows of 9)."""
    rxyz = np.asarray(rxyz, float)
    eps = np.asarray(eps, float)
    R = np.asarray(R, float).reshape(-1, 3, 3)
    d = np.asarray(d, float)
    n = np.einsum("kji,kj->ki", R, d)  # local direction = R^T d
    rx, ry, rz = rxyz[:, 0], rxyz[:, 1], rxyz[:, 2]
    e1, e2 = eps[:, 0], eps[:, 1]
    a1, a2 = 1.0 / (2.0 - e1), 1.0 / (2.0 - e2)
    ph0 = np.arctan2(np.sign(n[:, 1]) * np.abs(ry * n[:, 1]) ** a1, np.sign(n[:, 0]) * np.abs(rx * n[:, 0]) ** a1)
    ph1 = np.arctan2(np.sign(n[:, 2]) * np.abs(rz * n[:, 2] * np.abs(np.cos(ph0)) ** (2.0 - e1)) ** a2,
                     np.abs(rx * n[:, 0]) ** a2)
    c0, s0, c1, s1 = np.cos(ph0), np.sin(ph0), np.cos(ph1), np.sin(ph1)
    p = np.stack([np.sign(c0) * rx * np.abs(c0) ** e1 * np.abs(c1) ** e2,
                  np.sign(s0) * ry * np.abs(s0) ** e1 * np.abs(c1) ** e2,
                  np.sign(s1) * rz * np.abs(s1) ** e2], axis=1)
    return np.einsum("ki,ki->k", p, n)


def _empty_arrays(n):
    A = {
        "kind": np.zeros(n, np.int32), "pos": np.zeros((n, 3)), "ori": np.tile([1.0, 0, 0, 0], (n, 1)),
        "vel": np.zeros((n, 3)), "angVel": np.zeros((n, 3)), "angMom": np.zeros((n, 3)), "mass": np.zeros(n),
        "inertia": np.zeros((n, 3)), "rxyz": np.zeros((n, 3)), "eps": np.ones((n, 2)), "r_max": np.zeros(n),
        "isSphere": np.zeros(n, np.int32), "blockedDOFs": np.zeros(n, np.int32), "isDamped": np.ones(n, np.int32),
        "material": np.zeros(n, np.int32), "wallAxis": np.zeros(n, np.int32), "wallSense": np.zeros(n, np.int32),
        "hasBound": np.zeros(n, np.int32),
    }
    assert set(A) == set(BODY_ARRAYS)
    return A


def _box_walls(A, lo, hi, wall_mat):
    """utils.wall for the six faces, sense pointing inwards, ids 0..5."""
    k = 0
    for axis in range(3):
        for pos, sense in ((lo[axis], 1), (hi[axis], -1)):
            A["kind"][k] = nat.SDM_BODY_WALL
            A["pos"][k, axis] = pos
            A["wallAxis"][k] = axis
            A["wallSense"][k] = sense
            A["blockedDOFs"][k] = DOF_ALL
            A["material"][k] = wall_mat
            A["hasBound"][k] = 0
            k += 1


def _fill_particles(A, start, pos, rxyz, eps, quat, mat_index, density):
    n = len(pos)
    sl = slice(start, start + n)
    rmax, vol, inert = superquadric_props(rxyz, eps)
    A["kind"][sl] = nat.SDM_BODY_SUPERQUADRIC
    A["pos"][sl] = pos
    A["ori"][sl] = quat
    A["rxyz"][sl] = rxyz
    A["eps"][sl] = eps
    A["r_max"][sl] = rmax
    A["mass"][sl] = density * vol
    A["inertia"][sl] = inert * density
    A["material"][sl] = mat_index
    A["hasBound"][sl] = 1
    return vol


class Config:
    """A synthetic scene: body arrays, materials and the engine list."""

    def __init__(self, name, arrays, materials, engines, dt, steps, box, cell=None):
        self.name, self.arrays, self.materials, self.engines = name, arrays, materials, engines
        self.dt, self.steps, self.box = dt, steps, box
        self.cell = cell  # periodic cell size, or None

    @property
    def n(self):
        return len(self.arrays["kind"])

    def params(self, precision=nat.SDM_FP64):
        return engines_to_params(self.engines, self.dt, precision, self.cell)

    def open(self, lib=None, precision=nat.SDM_FP64):
        from .scene import open_arrays
        return open_arrays(lib or nat.load_hip(), self.params(precision), self.arrays, self.materials)

    def update_stats(self):
        """solids_fraction: particle volume over the box footprint times the bed height (top of
        the highest bounding sphere); mean_scale: mean semi-axis scale against the drawn sizes."""
        A = self.arrays
        part = A["kind"] == nat.SDM_BODY_SUPERQUADRIC
        if self.cell is not None:  # periodic: particle volume over the cell volume
            vol = float((A["mass"][part] / self.materials[A["material"][part][0]].density).sum())
            self.solids_fraction = vol / float(np.prod(self.cell))
            drawn = getattr(self, "rxyz_drawn", None)
            self.mean_scale = float((A["rxyz"][part, 0] / drawn[part, 0]).mean()) if drawn is not None else 1.0
            return
        lo, hi = self.box
        top = float((A["pos"][part, 2] + A["r_max"][part]).max()) - lo[2]
        vol = float((A["mass"][part] / self.materials[A["material"][part][0]].density).sum())
        self.solids_fraction = vol / ((hi[0] - lo[0]) * (hi[1] - lo[1]) * top)
        drawn = getattr(self, "rxyz_drawn", None)
        self.mean_scale = float((A["rxyz"][part, 0] / drawn[part, 0]).mean()) if drawn is not None else 1.0


def linear_engines(verlet):
    return [ForceResetter(), InsertionSortCollider([Bo1_Superquadrics_Aabb(), Bo1_Wall_Aabb()], verletDist=verlet),
            InteractionLoop([Ig2_Wall_Superquadrics_SuperquadricsGeom(),
                             Ig2_Superquadrics_Superquadrics_SuperquadricsGeom()],
                            [Ip2_SuperquadricsMat_SuperquadricsMat_SuperquadricsPhys()], [SuperquadricsLaw()]),
            NewtonIntegrator(damping=0.3, gravity=(0.0, 0.0, -9.81), isSuperquadrics=1)]


def hertz_engines(verlet):
    return [ForceResetter(), InsertionSortCollider([Bo1_Superquadrics_Aabb(), Bo1_Wall_Aabb()], verletDist=verlet),
            InteractionLoop([Ig2_Wall_Superquadrics_SuperquadricsGeom2(),
                             Ig2_Superquadrics_Superquadrics_SuperquadricsGeom2()],
                            [Ip2_SuperquadricsMat2_SuperquadricsMat2_HertzMindlinPhys()], [SuperquadricsLaw2()]),
            NewtonIntegrator(damping=0.3, gravity=(0.0, 0.0, -9.81), isSuperquadrics=1)]


def _draw_shapes(rng, n):
    rxyz = rng.uniform(0.8e-3, 1.2e-3, size=(n, 3))
    eps = rng.uniform(0.6, 1.4, size=(n, 2))
    return rxyz, eps


def _neighbour_pairs(pos, cutoff):
    """All pairs (i < j) of points closer than cutoff."""
    from scipy.spatial import cKDTree
    pr = cKDTree(pos).query_pairs(cutoff, output_type="ndarray")
    if len(pr) == 0:
        return np.zeros(0, int), np.zeros(0, int)
    pr = pr[np.lexsort((pr[:, 1], pr[:, 0]))]
    return pr[:, 0].astype(np.int64), pr[:, 1].astype(np.int64)


def _centre_line_overlap(pos, rxyz, eps, R, i, j):
    d = pos[j] - pos[i]
    L = np.linalg.norm(d, axis=1)
    u = d / L[:, None]
    return support_h(rxyz[i], eps[i], R[i], u) + support_h(rxyz[j], eps[j], R[j], -u) - L


def config0(n_particles=1000, seed_extra=0):
    """1,000 superellipsoids in a 20x20x60 mm box, linear law, FP64 deposition.

    Centres: the 7x7x21 lattice of spacing 2.86 mm (1,029 sites, 1,000 drawn), jittered by
    up to 0.1 mm.  A particle whose support along the line of centres overlaps a lattice
    neighbour (or a wall) gets a new orientation, up to 200 draws, then its semi-axes are
    scaled down until it is clear."""
    rng = rng_for(0, seed_extra)
    box_lo = np.zeros(3)
    box_hi = np.array([20e-3, 20e-3, 60e-3])
    sp = 2.86e-3
    grid = np.array([[i, j, k] for k in range(21) for j in range(7) for i in range(7)], float)
    sites = (grid + 0.5) * sp
    sites[:, 2] += 0.0
    pick = np.sort(rng.choice(len(sites), size=n_particles, replace=False))
    pos = sites[pick] + rng.uniform(-0.1e-3, 0.1e-3, size=(n_particles, 3))
    rxyz, eps = _draw_shapes(rng, n_particles)
    quat = random_quaternions(rng, n_particles)
    _separate(rng, pos, rxyz, eps, quat, box_lo, box_hi, gap=1e-6)
    p_mat = SuperquadricsMat(label="mat1", Kn=1e5, Ks=7e4, frictionAngle=math.atan(0.5), density=2650.0)
    w_mat = SuperquadricsMat(label="wallmat", Kn=1e5, Ks=7e4, frictionAngle=0.0, density=2650.0)
    A = _empty_arrays(6 + n_particles)
    _box_walls(A, box_lo, box_hi, wall_mat=1)
    _fill_particles(A, 6, pos, rxyz, eps, quat, 0, p_mat.density)
    verlet = 0.1 * float(np.mean(A["r_max"][6:]))
    return Config("C0: 1000 superellipsoids, linear law, 20x20x60 mm box", A, [p_mat, w_mat], linear_engines(verlet),
                  dt=1e-6, steps=20000, box=(box_lo, box_hi))


def _separate(rng, pos, rxyz, eps, quat, lo, hi, gap):
    """Re-draw orientations (then shrink) until no centre-line overlap with neighbours or walls."""
    n = len(pos)
    cutoff = 2 * 1.2e-3 * math.sqrt(3) + 1e-6
    ii, jj = _neighbour_pairs(pos, cutoff)
    for sweep in range(400):
        R = rot_matrices(quat)
        bad = np.zeros(n, bool)
        if len(ii):
            ov = _centre_line_overlap(pos, rxyz, eps, R, ii, jj)
            hit = ov > -gap
            bad[jj[hit]] = True
        for axis in range(3):
            e = np.zeros((n, 3))
            e[:, axis] = 1.0
            hp = support_h(rxyz, eps, R, e)
            hm = support_h(rxyz, eps, R, -e)
            bad |= (pos[:, axis] + hp > hi[axis] - gap) | (pos[:, axis] - hm < lo[axis] + gap)
        if not bad.any():
            return
        k = np.nonzero(bad)[0]
        if sweep < 200:
            quat[k] = random_quaternions(rng, len(k))
        else:
            rxyz[k] *= 0.97
    raise RuntimeError("could not separate the initial packing")


def config1(n_particles=200000, seed_extra=0, spacing=2.0e-3, box_xy=120e-3):
    """200,000 superellipsoids in a 120x120x200 mm box, Hertz-Mindlin law: the initial lattice.

    Centres on a face-centred cubic lattice of nearest-neighbour distance 2.0 mm filling the
    bottom of the box (jittered by up to 0.02 mm); semi-axes U[0.8,1.2] mm and exponents
    U[0.6,1.4] drawn per particle, then scaled by the largest factor <= 1 that keeps the
    particle's support along every line of centres to a lattice neighbour, and along the
    wall normals, inside the gap: no two particles overlap, so the lattice has no contacts.
    `densify` turns it into a contact-carrying bed.  The drawn (unscaled) semi-axes are kept
    as `cfg.rxyz_drawn`."""
    rng = rng_for(1, seed_extra)
    box_lo = np.zeros(3)
    box_hi = np.array([box_xy, box_xy, 200e-3])
    a = spacing
    cc = a * math.sqrt(2.0)
    nx = int((box_hi[0] - 1.6e-3) / cc)
    ny = int((box_hi[1] - 1.6e-3) / cc)
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    per_layer = nx * ny * 4
    nz = int(math.ceil(n_particles / per_layer))
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), axis=-1).reshape(-1, 3)
    sites = (g[:, None, :] + basis[None, :, :]).reshape(-1, 3) * cc
    sites = sites[np.lexsort((sites[:, 0], sites[:, 1], sites[:, 2]))][:n_particles]
    sites += 1.6e-3
    pos = sites + rng.uniform(-0.02e-3, 0.02e-3, size=sites.shape)
    rxyz, eps = _draw_shapes(rng, n_particles)
    drawn = rxyz.copy()
    quat = random_quaternions(rng, n_particles)
    R = rot_matrices(quat)
    ii, jj = _neighbour_pairs(pos, a * 1.5)
    scale = np.ones(n_particles)
    # pairwise bound: lambda_i, lambda_j <= L / (h_i + h_j)
    d = pos[jj] - pos[ii]
    L = np.linalg.norm(d, axis=1)
    u = d / L[:, None]
    hs = support_h(rxyz[ii], eps[ii], R[ii], u) + support_h(rxyz[jj], eps[jj], R[jj], -u)
    lim = np.minimum(1.0, L / hs)
    np.minimum.at(scale, ii, lim)
    np.minimum.at(scale, jj, lim)
    for axis in range(3):
        e = np.zeros((n_particles, 3))
        e[:, axis] = 1.0
        hp = support_h(rxyz, eps, R, e)
        hm = support_h(rxyz, eps, R, -e)
        scale = np.minimum(scale, np.minimum((box_hi[axis] - pos[:, axis]) / hp, (pos[:, axis] - box_lo[axis]) / hm))
    rxyz = rxyz * scale[:, None]
    p_mat = SuperquadricsMat2(label="sq", G=7e7 / (2 * (1 + 0.3)), nu=0.3, frictionAngle=math.atan(0.5),
                              density=2650.0, Kn=1e5, Ks=7e4)
    w_mat = SuperquadricsMat2(label="wall", G=7e7 / (2 * (1 + 0.3)), nu=0.3, frictionAngle=math.atan(0.5),
                              density=2650.0, Kn=1e5, Ks=7e4)
    A = _empty_arrays(6 + n_particles)
    _box_walls(A, box_lo, box_hi, wall_mat=1)
    _fill_particles(A, 6, pos, rxyz, eps, quat, 0, p_mat.density)
    verlet = 0.1 * float(np.mean(A["r_max"][6:]))
    cfg = Config(f"C1: {n_particles} superellipsoids, Hertz-Mindlin law, {box_xy * 1e3:.0f}x{box_xy * 1e3:.0f}x200 mm box",
                 A, [p_mat, w_mat], hertz_engines(verlet), dt=1e-6, steps=5000, box=(box_lo, box_hi))
    cfg.rxyz_drawn = np.concatenate([np.zeros((6, 3)), drawn])
    cfg.update_stats()
    return cfg


def densify(cfg, lib, iterations=150, target=2e-6, growth=0.003, max_push=5e-5, relax=25, log=None):
    """Grow-and-relax densification of a lattice configuration (a force-biased packing
    algorithm) evaluated by the contact engine itself.

    Each iteration runs one engine step with dt = 0 (geometry only: Ig2 gives every
    interaction's penetration depth and normal), then (1) moves both particles of every
    contact deeper than `target` apart along the contact normal by 0.35 x the excess (capped at
    `max_push`; a wall contact moves the particle by 0.7 x the excess), and (2) grows every
    particle whose deepest contact is shallower than 2 x `target` by the factor 1 + `growth`,
    up to its drawn semi-axes; the last `relax` iterations only move.  The collider is re-run
    every iteration.  Masses, inertia and
    r_max are recomputed from the final semi-axes (Superquadrics::Initial); velocities stay
    zero.  Deterministic for a given engine library."""
    A = cfg.arrays
    n = cfg.n
    ctx = cfg.open(lib)
    try:
        ctx.set("scene.dt", np.array([0.0]))
        part = A["kind"] == nat.SDM_BODY_SUPERQUADRIC
        rx0 = A["rxyz"].copy()
        rm0 = A["r_max"].copy()
        with np.errstate(divide="ignore", invalid="ignore"):
            cap = np.where(part, cfg.rxyz_drawn[:, 0] / np.where(part, rx0[:, 0], 1.0), 1.0)
        lam = np.ones(n)
        pos = A["pos"].copy()
        for it in range(iterations):
            if it:
                ctx.set("shape.rxyz", rx0 * lam[:, None])
                ctx.set("shape.r_max", rm0 * lam)
                ctx.set("state.pos", pos)
            ctx.set("collider.strideActive", np.array([0], np.int32))
            ctx.step(1)
            i1 = ctx.get("intr.id1")
            i2 = ctx.get("intr.id2")
            # a Levenberg-Marquardt solve that ends non-finite after its restarts leaves a NaN depth
            # (the reference stores it too); such a pair pushes nothing
            d = np.nan_to_num(ctx.get("geom.PenetrationDepth"), nan=0.0)
            nrm = np.nan_to_num(ctx.get("geom.normal").reshape(-1, 3), nan=0.0)
            ex = np.minimum(np.where(d > target, d - target, 0.0), max_push)
            w = np.where(A["kind"][i1] == nat.SDM_BODY_WALL, 0.7, 0.35) * ex
            for k in range(3):
                pos[:, k] += (np.bincount(i2, w * nrm[:, k], minlength=n) - np.bincount(i1, w * nrm[:, k], minlength=n)) * part
            dm = np.zeros(n)
            np.maximum.at(dm, i1, d)
            np.maximum.at(dm, i2, d)
            gr = part & (dm < 2 * target) & (lam < cap) & (it < iterations - relax)
            lam[gr] = np.minimum(cap[gr], lam[gr] * (1 + growth))
            if log and (it % 10 == 0 or it == iterations - 1):
                log("densify %d: %d contacts, %d growing" % (it, int((d > 0).sum()), int(gr.sum())))
    finally:
        ctx.close()
    B = {k: v.copy() for k, v in A.items()}
    idx = np.nonzero(part)[0]
    rxyz = rx0[idx] * lam[idx, None]
    B["pos"][idx] = pos[idx]
    B["rxyz"][idx] = rxyz
    rmax, vol, inert = superquadric_props(rxyz, B["eps"][idx])
    rho = cfg.materials[0].density
    B["r_max"][idx] = rmax
    B["mass"][idx] = rho * vol
    B["inertia"][idx] = inert * rho
    out = Config(cfg.name + ", densified", B, cfg.materials, copy.deepcopy(cfg.engines), cfg.dt, cfg.steps, cfg.box,
                 cfg.cell)
    out.rxyz_drawn = cfg.rxyz_drawn
    out.engines[1].verletDist = 0.1 * float(np.mean(B["r_max"][idx]))
    out.update_stats()
    return out


def box_fraction(cfg):
    """Solids fraction as CompressionEngine computes it (pkg/dem/Compression.cpp:425-458):
    the particles' volume over the volume of the box the six walls enclose."""
    A = cfg.arrays
    part = A["kind"] == nat.SDM_BODY_SUPERQUADRIC
    vol = float((A["mass"][part] / np.array([cfg.materials[k].density for k in A["material"][part]])).sum())
    lo, hi = _wall_box(A)
    return vol / float(np.prod(hi - lo))


def _wall_box(A):
    """Lower and upper box corners from the six walls (sense +1: the lower face of its axis)."""
    lo, hi = np.zeros(3), np.zeros(3)
    for k in np.nonzero(A["kind"] == nat.SDM_BODY_WALL)[0]:
        a = A["wallAxis"][k]
        (lo if A["wallSense"][k] > 0 else hi)[a] = A["pos"][k, a]
    return lo, hi


def config2(n_particles=1000000, seed_extra=0, fraction=0.62):
    """Periodic workload: n_particles superellipsoids in a cubic periodic cell, linear law.

    Shapes: rx = ry = r with r lognormal of mean 1 mm and log-sigma 0.25, rz = U[0.5, 1.5] r,
    eps1, eps2 U[0.3, 1.9].  The cell is sized so that the drawn particles fill `fraction` of
    it.  Centres: a face-centred cubic lattice over the cell (n sites drawn from its sites),
    jittered by 2% of the spacing; every particle is then scaled by the largest factor <= 1
    that keeps its support along each line of centres to a lattice neighbour (nearest periodic
    image) inside the gap, so the lattice has no contacts; `densify` grows it towards the
    drawn sizes.  No walls, no gravity.  The drawn semi-axes are kept as `cfg.rxyz_drawn`."""
    from scipy.spatial import cKDTree
    rng = rng_for(2, seed_extra)
    n = int(n_particles)
    sig = 0.25
    r = rng.lognormal(mean=math.log(1e-3) - 0.5 * sig * sig, sigma=sig, size=n)
    rxyz = np.stack([r, r, r * rng.uniform(0.5, 1.5, n)], axis=1)
    eps = rng.uniform(0.3, 1.9, size=(n, 2))
    _, vol, _ = superquadric_props(rxyz, eps)
    L = float((vol.sum() / fraction) ** (1.0 / 3.0))
    m = int(math.ceil((n / 4.0) ** (1.0 / 3.0)))
    cc = L / m
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    g = np.stack(np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij"), axis=-1).reshape(-1, 3)
    sites = ((g[:, None, :] + basis[None, :, :]).reshape(-1, 3) + 0.25) * cc
    pick = np.sort(rng.choice(len(sites), size=n, replace=False))
    a = cc / math.sqrt(2.0)  # nearest-neighbour distance
    pos = sites[pick] + rng.uniform(-0.02 * a, 0.02 * a, size=(n, 3))
    quat = random_quaternions(rng, n)
    R = rot_matrices(quat)
    tree = cKDTree(np.mod(pos, L), boxsize=L)
    pairs = tree.query_pairs(a * 1.5, output_type="ndarray")
    ii, jj = pairs[:, 0], pairs[:, 1]
    d = pos[jj] - pos[ii]
    d -= L * np.round(d / L)
    Ld = np.linalg.norm(d, axis=1)
    u = d / Ld[:, None]
    hs = support_h(rxyz[ii], eps[ii], R[ii], u) + support_h(rxyz[jj], eps[jj], R[jj], -u)
    lim = np.minimum(1.0, (Ld - 1e-7) / hs)
    scale = np.ones(n)
    np.minimum.at(scale, ii, lim)
    np.minimum.at(scale, jj, lim)
    drawn = rxyz.copy()
    rxyz = rxyz * scale[:, None]
    p_mat = SuperquadricsMat(label="sq", Kn=1e5, Ks=7e4, frictionAngle=math.atan(0.5), density=2650.0)
    A = _empty_arrays(n)
    _fill_particles(A, 0, pos, rxyz, eps, quat, 0, p_mat.density)
    verlet = 0.1 * float(np.mean(A["r_max"]))
    engines = linear_engines(verlet)
    engines[1].boundFunctors = [Bo1_Superquadrics_Aabb()]
    engines[2].geomFunctors = [Ig2_Superquadrics_Superquadrics_SuperquadricsGeom()]
    engines[3].gravity = (0.0, 0.0, 0.0)
    cfg = Config(f"C2: {n} superellipsoids, eps in [0.3,1.9], periodic cell {L * 1e3:.1f} mm, linear law", A, [p_mat],
                 engines, dt=1e-6, steps=1000, box=(np.zeros(3), np.full(3, L)), cell=(L, L, L))
    cfg.rxyz_drawn = drawn
    cfg.update_stats()
    return cfg


def compress(cfg, lib, target=0.62, growth=1e-3, every=50, max_steps=8000, max_depth=1e-5, relax=200, damping=0.7,
             log=None):
    """Growth under dynamics for a periodic packing: the engine runs with its own dt, local
    damping `damping` and frictionless copies of the materials; every `every` steps all
    particles grow by the factor 1 + `growth` (semi-axes, r_max, mass and inertia scaled as
    size^1, ^1, ^3, ^5) and all velocities are zeroed, until the solids fraction reaches `target`, the mean contact overlap exceeds
    `max_depth` in the median (jammed), or `max_steps`; then `relax` steps without growth and all
    velocities zeroed.  Returns a new Config with the final positions, orientations and sizes
    (the original materials)."""
    from .scene import SuperquadricsMat as _M
    assert cfg.cell is not None, "compress() is for periodic cells"
    A = cfg.arrays
    n = cfg.n
    mats = []
    for m in cfg.materials:
        c = copy.copy(m)
        c.frictionAngle = 0.0
        mats.append(c)
    eng = copy.deepcopy(cfg.engines)
    eng[3].damping = damping
    ctx = Config(cfg.name, A, mats, eng, cfg.dt, cfg.steps, cfg.box, cfg.cell).open(lib)
    zero = np.zeros((n, 3))
    vol = A["mass"] / np.array([cfg.materials[k].density for k in A["material"]])
    cellvol = float(np.prod(cfg.cell))
    scale, steps, frac = 1.0, 0, float(vol.sum()) / cellvol
    try:
        while steps < max_steps and frac < target:
            ctx.step(every)
            steps += every
            g = 1.0 + growth
            scale *= g
            ctx.set("shape.rxyz", A["rxyz"] * scale)
            ctx.set("shape.r_max", A["r_max"] * scale)
            ctx.set("state.mass", A["mass"] * scale ** 3)
            ctx.set("state.inertia", A["inertia"] * scale ** 5)
            for k in ("state.vel", "state.angVel", "state.angMom"):  # quench: a slow, cold compression
                ctx.set(k, zero)
            frac = float(vol.sum()) * scale ** 3 / cellvol
            if steps % (10 * every) == 0:
                d = np.nan_to_num(ctx.get("geom.PenetrationDepth"), nan=0.0)
                md = float(np.median(d[d > 0])) if (d > 0).any() else 0.0
                if log:
                    log("compress %d steps: solids fraction %.3f, %d contacts, median overlap %.2e" % (
                        steps, frac, int((d > 0).sum()), md))
                if md > max_depth:
                    break
        ctx.step(relax)
        B = {k: v.copy() for k, v in A.items()}
        B["pos"] = ctx.get("state.pos").reshape(-1, 3).copy()
        B["ori"] = ctx.get("state.ori").reshape(-1, 4).copy()
        B["rxyz"] = A["rxyz"] * scale
        B["r_max"] = A["r_max"] * scale
        B["mass"] = A["mass"] * scale ** 3
        B["inertia"] = A["inertia"] * scale ** 5
        for k in ("vel", "angVel", "angMom"):
            B[k] = np.zeros_like(A[k])
    finally:
        ctx.close()
    out = Config(cfg.name + ", compressed", B, cfg.materials, copy.deepcopy(cfg.engines), cfg.dt, cfg.steps, cfg.box,
                 cfg.cell)
    out.rxyz_drawn = getattr(cfg, "rxyz_drawn", None)
    out.update_stats()
    return out


def tile(cfg, k):
    """k x k x k copies of a periodic configuration (positions wrapped into the cell first) in a
    cell k times larger: the same packing, solids fraction and contacts, k^3 times the bodies."""
    assert cfg.cell is not None
    L0 = np.array(cfg.cell)
    A = cfg.arrays
    pos = np.mod(A["pos"], L0)
    offs = np.array([[i, j, l] for l in range(k) for j in range(k) for i in range(k)], float) * L0
    B = {key: np.concatenate([v] * len(offs)) for key, v in A.items()}
    B["pos"] = np.concatenate([pos + o for o in offs])
    out = Config(cfg.name + ", tiled %dx%dx%d" % (k, k, k), B, cfg.materials, copy.deepcopy(cfg.engines), cfg.dt,
                 cfg.steps, (np.zeros(3), L0 * k), tuple(float(x) for x in L0 * k))
    drawn = getattr(cfg, "rxyz_drawn", None)
    out.rxyz_drawn = np.concatenate([drawn] * len(offs)) if drawn is not None else None
    out.update_stats()
    return out


def config2_dense(lib, n_particles=1000000, base=8000, log=None):
    """The periodic workload as a dense packing: a `base`-particle cell of config2's
    distributions, densified (grow-and-relax) and compressed under dynamics to ~0.6 solids
    fraction, then tiled k x k x k (k^3 base = n_particles) -- compressing a million bodies
    directly would take tens of minutes of engine time."""
    k = int(round((n_particles / float(base)) ** (1.0 / 3.0)))
    assert k >= 1 and k ** 3 * base == n_particles, "n_particles must be k^3 x base"
    c = config2(n_particles=base)
    c = densify(c, lib, iterations=60, growth=0.01, relax=20)
    if log:
        log("base cell: %d particles densified to solids fraction %.3f" % (base, c.solids_fraction))
    c = compress(c, lib, log=log)
    if log:
        log("base cell compressed to solids fraction %.3f" % c.solids_fraction)
    out = tile(c, k) if k > 1 else c
    out.base = c
    return out


def box_specimen(n_particles, rng, kn, mu, spacing=2.86e-3):
    """A config-0-style specimen of any size: the 20x20x60 mm proportions of config0 (a
    lattice nx x nx x 3nx of 2.86 mm), n_particles drawn sites, jitter 0.1 mm, config-0 shape
    distributions, each particle scaled down to clear its lattice neighbours and the walls along
    the lines of centres (vectorised, as config1).  Linear law: particle Kn = kn, Ks = 0.7 kn,
    friction mu; walls frictionless with Kn 1e5, Ks 7e4 (config0)."""
    nx = max(2, int(math.ceil((n_particles / 3.0) ** (1.0 / 3.0))))
    nz = int(math.ceil(n_particles / float(nx * nx)))
    box_lo = np.zeros(3)
    box_hi = np.array([nx * spacing, nx * spacing, nz * spacing])
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(nx), np.arange(nz), indexing="ij"), axis=-1).reshape(-1, 3)
    pick = np.sort(rng.choice(len(g), size=n_particles, replace=False))
    pos = (g[pick] + 0.5) * spacing + rng.uniform(-0.1e-3, 0.1e-3, size=(n_particles, 3))
    rxyz, eps = _draw_shapes(rng, n_particles)
    quat = random_quaternions(rng, n_particles)
    R = rot_matrices(quat)
    scale = np.ones(n_particles)
    ii, jj = _neighbour_pairs(pos, spacing * 1.8)
    if len(ii):
        d = pos[jj] - pos[ii]
        L = np.linalg.norm(d, axis=1)
        u = d / L[:, None]
        hs = support_h(rxyz[ii], eps[ii], R[ii], u) + support_h(rxyz[jj], eps[jj], R[jj], -u)
        lim = np.minimum(1.0, (L - 1e-6) / hs)
        np.minimum.at(scale, ii, lim)
        np.minimum.at(scale, jj, lim)
    for axis in range(3):
        e = np.zeros((n_particles, 3))
        e[:, axis] = 1.0
        hp = support_h(rxyz, eps, R, e)
        hm = support_h(rxyz, eps, R, -e)
        scale = np.minimum(scale, np.minimum((box_hi[axis] - pos[:, axis] - 1e-6) / hp,
                                             (pos[:, axis] - box_lo[axis] - 1e-6) / hm))
    rxyz = rxyz * scale[:, None]
    p_mat = SuperquadricsMat(label="sq", Kn=kn, Ks=0.7 * kn, frictionAngle=math.atan(mu), density=2650.0)
    w_mat = SuperquadricsMat(label="wall", Kn=1e5, Ks=7e4, frictionAngle=0.0, density=2650.0)
    A = _empty_arrays(6 + n_particles)
    _box_walls(A, box_lo, box_hi, wall_mat=1)
    _fill_particles(A, 6, pos, rxyz, eps, quat, 0, p_mat.density)
    verlet = 0.1 * float(np.mean(A["r_max"][6:]))
    return Config(f"box specimen of {n_particles} superellipsoids", A, [p_mat, w_mat], linear_engines(verlet),
                  dt=1e-6, steps=10000, box=(box_lo, box_hi))


def batch(cfgs):
    """Several independent specimens in one context: bodies concatenated specimen by specimen
    (each with its six walls first), materials concatenated, `specimen` = the specimen index.
    All specimens share the engine list and dt of the first (verletDist: the smallest)."""
    n = sum(c.n for c in cfgs)
    A = {k: np.concatenate([c.arrays[k] for c in cfgs]) for k in cfgs[0].arrays}
    mats, spec, off = [], [], 0
    for k, c in enumerate(cfgs):
        A["material"][off:off + c.n] = c.arrays["material"] + len(mats)
        mats.extend(c.materials)
        spec.append(np.full(c.n, k, np.int32))
        off += c.n
    A["specimen"] = np.concatenate(spec)
    engines = copy.deepcopy(cfgs[0].engines)
    engines[1].verletDist = min(c.engines[1].verletDist for c in cfgs)
    out = Config(f"batch of {len(cfgs)} specimens, {n} bodies", A, mats, engines, cfgs[0].dt, cfgs[0].steps,
                 cfgs[0].box)
    out.sizes = [c.n for c in cfgs]
    return out


def config4(n_specimens=256, seed_extra=0, n_min=2000, n_max=50000):
    """Batching workload: n_specimens box specimens with particle counts log-uniform in
    [n_min, n_max], each with its own Kn ~ U[5e4, 2e5] and friction coefficient ~ U[0.2, 0.8]."""
    rng = rng_for(4, seed_extra)
    sizes = np.exp(rng.uniform(math.log(n_min), math.log(n_max), n_specimens)).round().astype(int)
    kns = rng.uniform(5e4, 2e5, n_specimens)
    mus = rng.uniform(0.2, 0.8, n_specimens)
    specs = [box_specimen(int(sizes[k]), rng, float(kns[k]), float(mus[k])) for k in range(n_specimens)]
    for sp in specs:  # one verletDist for all: 0.1 x the mean r_max of the whole batch
        sp.engines[1].verletDist = 0.1 * float(np.mean(np.concatenate([c.arrays["r_max"][6:] for c in specs])))
    return specs


TOP_WALL = 5  # the z-max wall of _box_walls (sense -1)


def config3(lib, n_particles=20000, box_xy=40e-3, wall_speed=0.05, settle_steps=20000, goal=2e4, log=None):
    """Differentiation workload: n_particles superellipsoids of the plain workload's size and
    shape distributions and linear law (Kn 1e5, Ks 7e4, mu 0.5, the walls too) on the
    lattice of config1 (box_xy wide), settled in the box by the wall servo (CompressionEngine,
    `goal` Pa on every wall, max_vel 1 m/s; gravity, local damping 0.3, dt 1e-6 s) for
    `settle_steps` steps, so that the packing is jammed and every wall carries load; the servo
    is then switched off and the top wall given the downward velocity `wall_speed` (Newton
    translates a blocked body with its velocity).  Returns the open engine context at that
    state (the caller freezes its contacts and closes it) and the Config."""
    cfg = config1(n_particles=n_particles, box_xy=box_xy)
    p_mat = SuperquadricsMat(label="mat1", Kn=1e5, Ks=7e4, frictionAngle=math.atan(0.5), density=2650.0)
    # the walls have the particles' friction, so that one (kn, ks, mu) describes every contact
    # of the frozen window, as the rollout's parameters do
    w_mat = SuperquadricsMat(label="wallmat", Kn=1e5, Ks=7e4, frictionAngle=math.atan(0.5), density=2650.0)
    cfg.materials = [p_mat, w_mat]
    cfg.engines = linear_engines(0.1 * float(np.mean(cfg.arrays["r_max"][6:])))
    A = cfg.arrays
    part = A["kind"] == nat.SDM_BODY_SUPERQUADRIC
    R = rot_matrices(A["ori"][part])
    for axis in range(3):  # walls on the lattice's outermost surfaces
        e = np.zeros((int(part.sum()), 3))
        e[:, axis] = 1.0
        A["pos"][2 * axis, axis] = float((A["pos"][part, axis] - support_h(A["rxyz"][part], A["eps"][part], R, -e)).min())
        A["pos"][2 * axis + 1, axis] = float((A["pos"][part, axis] + support_h(A["rxyz"][part], A["eps"][part], R, e)).max())
    ctx = cfg.open(lib)
    ctx.set_servo(goal=(goal, goal, goal), max_vel=1.0, gain_alpha=0.5, iterate_num_max=100, solve_num_max=2000,
                  echo_interval=1000)
    done = 0
    while done < settle_steps:
        k = min(5000, settle_steps - done)
        ctx.step(k)
        done += k
        if log:
            st = ctx.servo_state()
            log("settle %d: wall stresses %s" % (done, np.round(st["wall_stress"]).tolist()))
    ctx.set_servo(False)
    vel = ctx.get("state.vel").reshape(-1, 3)
    vel[TOP_WALL] = (0.0, 0.0, -wall_speed)
    ctx.set("state.vel", vel)
    return ctx, cfg
