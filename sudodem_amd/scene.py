"""Python layer mirroring the reference's scripting API for the superellipsoid step path.

A SudoDEM3D script builds a scene from materials, bodies (``NewSuperquadrics_rot2``,
``utils.wall``) and the engine list, then steps it (examples/SudoDEM3D/example1.py:76-120).
The same script runs here against the MI355X library::

    O = Omega()
    mat = SuperquadricsMat(Kn=1e5, Ks=7e4, frictionAngle=math.atan(0.5), density=2650)
    O.materials.append(mat)
    O.bodies.append(utils.wall(0.0, axis=2, sense=1, material=mat))
    b = NewSuperquadrics_rot2(1e-3, 1e-3, 1e-3, 1.0, 1.0, mat, 1, 0, 0, 0, False)
    b.state.pos = (0, 0, 1e-3); O.bodies.append(b)
    O.engines = [ForceResetter(),
                 InsertionSortCollider([Bo1_Superquadrics_Aabb(), Bo1_Wall_Aabb()], verletDist=1e-4),
                 InteractionLoop([Ig2_Wall_Superquadrics_SuperquadricsGeom(),
                                  Ig2_Superquadrics_Superquadrics_SuperquadricsGeom()],
                                 [Ip2_SuperquadricsMat_SuperquadricsMat_SuperquadricsPhys()],
                                 [SuperquadricsLaw()]),
                 NewtonIntegrator(damping=0.3, gravity=(0, 0, -9.81), isSuperquadrics=1)]
    O.dt = 1e-6
    O.run(100, True)

Only this engine list is accepted (anything else raises): it is the path the library
implements, optionally with a ``CompressionEngine`` (the cubic specimen's wall servo) between
the InteractionLoop and the NewtonIntegrator, as examples/SudoDEM3D/example2_consol.py has it.  Body data is held host-side until the first step, then lives on the device;
``O.bodies[i].state.pos`` and the bulk accessors ``O.get/O.set`` read or write it there.
"""
from __future__ import annotations

import math

import numpy as np

from . import _native as nat

DOF_ALL = 63


# ---------------------------------------------------------------------------------------
# Materials (pkg/dem/Superquadrics.hpp:247-277)
# ---------------------------------------------------------------------------------------
class SuperquadricsMat:
    """Elastic material with Coulomb friction (linear law)."""

    def __init__(self, label="", Kn=1e8, Ks=1e5, frictionAngle=0.5, betan=0.0, betas=0.0, density=1000.0, **kw):
        self.label = label
        self.Kn, self.Ks, self.frictionAngle = float(Kn), float(Ks), float(frictionAngle)
        self.betan, self.betas, self.density = float(betan), float(betas), float(density)
        self.G, self.nu = 29e9, 0.15
        for k, v in kw.items():
            setattr(self, k, v)
        self.id = -1

    def _native(self):
        return nat.sdm_material(self.density, self.Kn, self.Ks, self.frictionAngle, self.betan, self.betas, self.G,
                                self.nu)


class SuperquadricsMat2(SuperquadricsMat):
    """Material of the Hertz-Mindlin law: shear modulus G and Poisson's ratio nu."""

    def __init__(self, label="", G=29e9, nu=0.15, **kw):
        super().__init__(label=label, **kw)
        self.G, self.nu = float(G), float(nu)


# ---------------------------------------------------------------------------------------
# Bodies
# ---------------------------------------------------------------------------------------
class State:
    """core/State.hpp: position, orientation (w, x, y, z), velocities, mass, inertia."""

    def __init__(self):
        self.pos = np.zeros(3)
        self.ori = np.array([1.0, 0.0, 0.0, 0.0])
        self.vel = np.zeros(3)
        self.angVel = np.zeros(3)
        self.angMom = np.zeros(3)
        self.mass = 0.0
        self.inertia = np.zeros(3)
        self.blockedDOFs = 0
        self.isDamped = True


class Superquadrics:
    """Superellipsoid with semi-axes (rx, ry, rz) and exponents (eps1, eps2)."""

    def __init__(self, rx, ry, rz, eps1, eps2, isSphere=False):
        self.rxyz = np.array([rx, ry, rz], dtype=float)
        self.eps = np.array([eps1, eps2], dtype=float)
        self.isSphere = bool(isSphere)
        r, v, inert = nat.load_hip().props(rx, ry, rz, eps1, eps2)
        self.r_max, self.volume, self.inertia = r, v, inert

    def getVolume(self):
        return self.volume

    def getInertia(self):
        return self.inertia


class Wall:
    """Axis-aligned infinite plane; sense -1, 0 or +1 (pkg/common/Wall.hpp:8-28)."""

    def __init__(self, axis=0, sense=0):
        self.axis, self.sense = int(axis), int(sense)


class Body:
    def __init__(self, shape, material, state=None, bounded=False):
        self.shape = shape
        self.material = material
        self.state = state or State()
        self.hasBound = bounded  # NewSuperquadrics* allocate an Aabb; utils.wall does not
        self.id = -1
        self._scene = None


def _quat_to_rot(q):
    w, x, y, z = q
    tx, ty, tz = 2 * x, 2 * y, 2 * z
    twx, twy, twz = tx * w, ty * w, tz * w
    txx, txy, txz = tx * x, ty * x, tz * x
    tyy, tyz, tzz = ty * y, tz * y, tz * z
    return np.array([1 - (tyy + tzz), txy - twz, txz + twy, txy + twz, 1 - (txx + tzz), tyz - twx, txz - twy,
                     tyz + twx, 1 - (txx + tyy)])


def _quat_normalized(q):
    w, x, y, z = (float(v) for v in q)
    s2 = (x * x + y * y) + (z * z + w * w)
    if s2 > 0:
        s = math.sqrt(s2)
        return np.array([w / s, x / s, y / s, z / s])
    return np.array([w, x, y, z])


def NewSuperquadrics_rot2(x, y, z, ep1, ep2, mat, q0, q1, q2, q3, isSphere=False):
    """py/_superquadrics_utils.cpp:354-379: body with explicit orientation quaternion (w=q0)."""
    shape = Superquadrics(x, y, z, ep1, ep2, isSphere)
    b = Body(shape, mat, bounded=True)
    b.state.mass = mat.density * shape.volume
    b.state.inertia = shape.inertia * mat.density
    b.state.ori = _quat_normalized((q0, q1, q2, q3))
    return b


def NewSuperquadrics2(x, y, z, ep1, ep2, mat, rotate=False, isSphere=False):
    return NewSuperquadrics_rot2(x, y, z, ep1, ep2, mat, 1.0, 0.0, 0.0, 0.0, isSphere)


class utils:
    @staticmethod
    def wall(position, axis, sense=0, material=None, **kw):
        """py/utils.py:311-327: fixed wall body (all DOFs blocked, no bound until the collider)."""
        b = Body(Wall(axis, sense), material, bounded=False)
        p = np.zeros(3)
        if np.isscalar(position):
            p[axis] = float(position)
        else:
            p[:] = position
        b.state.pos = p
        b.state.blockedDOFs = DOF_ALL
        return b


# ---------------------------------------------------------------------------------------
# Engines and functors (names and attributes of the reference classes)
# ---------------------------------------------------------------------------------------
class _Functor:
    def __init__(self, **kw):
        self.__dict__.update(kw)


class Bo1_Superquadrics_Aabb(_Functor): pass
class Bo1_Wall_Aabb(_Functor): pass
class Ig2_Superquadrics_Superquadrics_SuperquadricsGeom(_Functor): pass
class Ig2_Superquadrics_Superquadrics_SuperquadricsGeom2(_Functor): pass
class Ig2_Wall_Superquadrics_SuperquadricsGeom(_Functor): pass
class Ig2_Wall_Superquadrics_SuperquadricsGeom2(_Functor): pass
class Ip2_SuperquadricsMat_SuperquadricsMat_SuperquadricsPhys(_Functor): pass
class Ip2_SuperquadricsMat2_SuperquadricsMat2_HertzMindlinPhys(_Functor): pass
class SuperquadricsLaw(_Functor): pass
class SuperquadricsLaw2(_Functor): pass


class ForceResetter(_Functor):
    pass


class InsertionSortCollider:
    def __init__(self, boundFunctors=(), verletDist=-0.5, targetInterv=50, minSweepDistFactor=0.1, sortAxis=0, **kw):
        self.boundFunctors = list(boundFunctors)
        self.verletDist, self.targetInterv = float(verletDist), int(targetInterv)
        self.minSweepDistFactor, self.sortAxis = float(minSweepDistFactor), int(sortAxis)
        self.__dict__.update(kw)


class InteractionLoop:
    def __init__(self, geomFunctors=(), physFunctors=(), lawFunctors=()):
        self.geomFunctors = list(geomFunctors)
        self.physFunctors = list(physFunctors)
        self.lawFunctors = list(lawFunctors)


class NewtonIntegrator:
    def __init__(self, damping=0.2, gravity=(0.0, 0.0, 0.0), isSuperquadrics=0, exactAsphericalRot=True, label="",
                 **kw):
        self.damping = float(damping)
        self.gravity = tuple(float(g) for g in gravity)
        self.isSuperquadrics = int(isSuperquadrics)
        self.exactAsphericalRot = bool(exactAsphericalRot)
        self.label = label
        self.__dict__.update(kw)


class CompressionEngine:
    """Wall stress servo of a cubic specimen (pkg/dem/Compression.hpp:160-210), placed between
    InteractionLoop and NewtonIntegrator (examples/SudoDEM3D/example2_consol.py:78-105).

    Attributes carry the reference's names and defaults.  Assigning one while the scene runs
    changes the running engine (``triax.z_servo = False`` starts shear with the state of the
    consolidation, as in example2_shear.py:60-71); ``dead`` switches it off and on.  The
    engine's state reads back as ``wsxx``, ``wsyy``, ``wszz``, ``width``, ``depth``,
    ``height``, ``UnbalancedForce`` and, after the first run, ``width0``, ``depth0``,
    ``height0``.  Assigning ``width0``/``depth0``/``height0`` takes effect at a first run with
    ``continueFlag``.  File output (``file``, ``savedata_interval``) and the console echo text
    are not produced."""

    _DEFAULTS = dict(goalx=0.0, goaly=0.0, goalz=0.0, max_vel=1.0, gain_alpha=0.5, f_threshold=0.01,
                     fmin_threshold=0.01, unbf_tol=0.01, iterate_num_max=100, solve_num_max=2000,
                     left_wall_activated=True, right_wall_activated=True, front_wall_activated=True,
                     back_wall_activated=True, bottom_wall_activated=True, top_wall_activated=True, z_servo=True,
                     hydroStrain=False, hydroStrainRate=0.0, shearMode=1, ramp_interval=10000, ramp_chunks=400,
                     echo_interval=1000, target_strain=0.15, continueFlag=False, width0=0.0, depth0=0.0,
                     height0=0.0, cubic_specimen=True, dead=False, label="", savedata_interval=12,
                     file="recordData", debug=False, interStressControl=False)
    _STATE = ("wsxx", "wsyy", "wszz", "width", "depth", "height", "UnbalancedForce")

    def __init__(self, **kw):
        object.__setattr__(self, "_p", dict(self._DEFAULTS))
        object.__setattr__(self, "_ctx", None)
        unknown = set(kw) - set(self._DEFAULTS)
        if unknown:
            raise AttributeError("CompressionEngine has no attribute(s) %s" % sorted(unknown))
        self._p.update(kw)

    def _servo_kw(self):
        p = self._p
        if not p["cubic_specimen"]:
            raise ValueError("only the cubic specimen is supported (cubic_specimen=True)")
        if p["interStressControl"]:
            raise ValueError("interStressControl is not supported (the SOLE_GAIN servo uses wall stresses)")
        return dict(goal=(p["goalx"], p["goaly"], p["goalz"]), max_vel=p["max_vel"], gain_alpha=p["gain_alpha"],
                    f_threshold=p["f_threshold"], fmin_threshold=p["fmin_threshold"], unbf_tol=p["unbf_tol"],
                    iterate_num_max=p["iterate_num_max"], solve_num_max=p["solve_num_max"],
                    activated=(p["left_wall_activated"], p["right_wall_activated"], p["front_wall_activated"],
                               p["back_wall_activated"]),
                    z_servo=p["z_servo"], hydro_strain=p["hydroStrain"], hydro_strain_rate=p["hydroStrainRate"],
                    shear_mode=p["shearMode"], ramp_interval=p["ramp_interval"], ramp_chunks=p["ramp_chunks"],
                    echo_interval=p["echo_interval"], target_strain=p["target_strain"],
                    continue_flag=p["continueFlag"], size0=(p["width0"], p["depth0"], p["height0"]))

    def _attach(self, ctx):
        object.__setattr__(self, "_ctx", ctx)
        object.__setattr__(self, "_set", False)
        self._push()

    def _push(self):
        ctx = self._ctx
        if ctx is None:
            return
        if self._p["dead"]:
            if self._set:
                ctx.set_servo(False)
            return
        ctx.set_servo(**self._servo_kw(), keep_state=1 if self._set else 0)
        object.__setattr__(self, "_set", True)

    def __getattr__(self, k):
        p = object.__getattribute__(self, "_p")
        ctx = object.__getattribute__(self, "_ctx")
        live = ctx is not None and self.__dict__.get("_set", False)
        if k in self._STATE or (live and k in ("width0", "depth0", "height0")):
            if not live:
                return 0.0 if k != "UnbalancedForce" else 1.0
            st = ctx.servo_state()
            if k == "UnbalancedForce":
                return float(st["unbalanced_force"])
            if k in ("wsxx", "wsyy", "wszz"):
                return float(st["stress"]["xyz".index(k[2])])
            if k in ("width", "depth", "height"):
                return float(st["size"][("width", "depth", "height").index(k)])
            return float(st["size0"][("width0", "depth0", "height0").index(k)])
        if k in p:
            return p[k]
        raise AttributeError(k)

    def __setattr__(self, k, v):
        if k in self._STATE:
            raise AttributeError("%s is read-only here (the engine's state)" % k)
        if k not in self._p:
            raise AttributeError("CompressionEngine has no attribute %s" % k)
        self._p[k] = v
        self._push()


class Cell:
    """Axis-aligned periodic cell (core/Cell.hpp): only its size is used (no velGrad)."""

    def __init__(self):
        self.size = (1.0, 1.0, 1.0)

    def setBox(self, x, y=None, z=None):
        if y is None:
            x, y, z = x
        self.size = (float(x), float(y), float(z))

    @property
    def hSize(self):
        return np.diag(self.size)


def engines_to_params(engines, dt, precision=nat.SDM_FP64, cell=None):
    """Validate the engine list and map its attributes onto sdm_params; `cell` (three sizes)
    makes the scene periodic (Scene::isPeriodic with an axis-aligned Cell)."""
    kinds = [type(e).__name__ for e in engines]
    if kinds not in (["ForceResetter", "InsertionSortCollider", "InteractionLoop", "NewtonIntegrator"],
                     ["ForceResetter", "InsertionSortCollider", "InteractionLoop", "CompressionEngine",
                      "NewtonIntegrator"]):
        raise ValueError("engine list must be [ForceResetter, InsertionSortCollider, InteractionLoop, "
                         f"(CompressionEngine,) NewtonIntegrator]; got {kinds}")
    col, loop, newton = engines[1], engines[2], engines[-1]
    if len(engines) == 5 and cell is not None:
        raise ValueError("CompressionEngine needs the six walls of a cubic specimen, not a periodic cell")
    bo = sorted(type(f).__name__ for f in col.boundFunctors)
    if bo and not set(bo) <= {"Bo1_Superquadrics_Aabb", "Bo1_Wall_Aabb"}:
        raise ValueError(f"unsupported bound functors {bo}")
    if col.verletDist < 0:
        raise ValueError("verletDist must be set explicitly: its automatic value needs spheres "
                         "(InsertionSortCollider.cpp:241-252)")
    geo = {type(f).__name__ for f in loop.geomFunctors}
    law = {type(f).__name__ for f in loop.lawFunctors}
    if law == {"SuperquadricsLaw"} and geo <= {"Ig2_Superquadrics_Superquadrics_SuperquadricsGeom",
                                              "Ig2_Wall_Superquadrics_SuperquadricsGeom"}:
        lawid = nat.SDM_LAW_LINEAR
    elif law == {"SuperquadricsLaw2"} and geo <= {"Ig2_Superquadrics_Superquadrics_SuperquadricsGeom2",
                                                 "Ig2_Wall_Superquadrics_SuperquadricsGeom2"}:
        lawid = nat.SDM_LAW_HERTZ
    else:
        raise ValueError(f"unsupported functor combination geom={sorted(geo)} law={sorted(law)}")
    if newton.isSuperquadrics != 1:
        raise ValueError("NewtonIntegrator.isSuperquadrics must be 1 for superellipsoids")
    p = nat.sdm_params()
    p.dt = float(dt)
    p.gravity[:] = newton.gravity
    p.damping = newton.damping
    p.exact_aspherical_rot = 1 if newton.exactAsphericalRot else 0
    p.verlet_dist = col.verletDist
    p.target_interv = col.targetInterv
    p.min_sweep_dist_factor = col.minSweepDistFactor
    p.sort_axis = col.sortAxis
    p.law = lawid
    p.precision = precision
    if cell is not None:
        p.periodic = 1
        p.cell[:] = [float(v) for v in cell]
    return p


# ---------------------------------------------------------------------------------------
# Bulk scene description (one row per body), used by the Omega front end and by the
# synthetic configurations.
# ---------------------------------------------------------------------------------------
BODY_ARRAYS = ("kind", "pos", "ori", "vel", "angVel", "angMom", "mass", "inertia", "rxyz", "eps", "r_max",
               "isSphere", "blockedDOFs", "isDamped", "material", "wallAxis", "wallSense", "hasBound")

_ARRAY_TO_ATTR = {
    "kind": "shape.kind", "pos": "state.pos", "ori": "state.ori", "vel": "state.vel", "angVel": "state.angVel",
    "angMom": "state.angMom", "mass": "state.mass", "inertia": "state.inertia", "rxyz": "shape.rxyz",
    "eps": "shape.eps", "r_max": "shape.r_max", "isSphere": "shape.isSphere", "blockedDOFs": "state.blockedDOFs",
    "isDamped": "state.isDamped", "material": "body.material", "wallAxis": "shape.wallAxis",
    "wallSense": "shape.wallSense", "hasBound": "bound.exists",
}


def upload_arrays(ctx: nat.Context, arrays: dict):
    """Write a bulk body description into an open context, with the rotation caches the
    body factories compute from the (normalised) orientation (Superquadrics.hpp:86-91)."""
    for k in BODY_ARRAYS:
        ctx.set(_ARRAY_TO_ATTR[k], arrays[k])
    ori = np.asarray(arrays["ori"], dtype=float).reshape(-1, 4)
    rg = rot_matrices(ori)
    conj = ori * np.array([1.0, -1.0, -1.0, -1.0])
    rl = rot_matrices(conj)
    ctx.set("shape.rot_mat2global", rg)
    ctx.set("shape.rot_mat2local", rl)
    if "specimen" in arrays:  # several independent specimens in one context (batching)
        ctx.set("body.specimen", np.asarray(arrays["specimen"], dtype=np.int32))


def rot_matrices(q):
    """Vectorised QuaternionBase::toRotationMatrix over rows (w, x, y, z)."""
    q = np.asarray(q, dtype=float).reshape(-1, 4)
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    tx, ty, tz = 2.0 * x, 2.0 * y, 2.0 * z
    twx, twy, twz = tx * w, ty * w, tz * w
    txx, txy, txz = tx * x, ty * x, tz * x
    tyy, tyz, tzz = ty * y, tz * y, tz * z
    R = np.empty((q.shape[0], 9))
    R[:, 0] = 1.0 - (tyy + tzz); R[:, 1] = txy - twz; R[:, 2] = txz + twy
    R[:, 3] = txy + twz; R[:, 4] = 1.0 - (txx + tzz); R[:, 5] = tyz - twx
    R[:, 6] = txz - twy; R[:, 7] = tyz + twx; R[:, 8] = 1.0 - (txx + tyy)
    return R


class Omega:
    """Scene + engine loop front end (py/wrapper/sudodemWrapper.cpp:929-991)."""

    def __init__(self, lib: nat.SdmLibrary | None = None, precision=nat.SDM_FP64):
        self.lib = lib
        self.precision = precision
        self.bodies = []
        self.materials = []
        self.engines = []
        self.dt = 1e-8
        self.ctx = None
        self.periodic = False  # Scene::isPeriodic
        self.cell = Cell()

    # -- lifecycle --
    def _arrays(self):
        mats = self.materials
        for i, m in enumerate(mats):
            m.id = i
        n = len(self.bodies)
        A = {k: None for k in BODY_ARRAYS}
        A["kind"] = np.zeros(n, np.int32)
        A["pos"] = np.zeros((n, 3)); A["ori"] = np.zeros((n, 4)); A["vel"] = np.zeros((n, 3))
        A["angVel"] = np.zeros((n, 3)); A["angMom"] = np.zeros((n, 3)); A["mass"] = np.zeros(n)
        A["inertia"] = np.zeros((n, 3)); A["rxyz"] = np.zeros((n, 3)); A["eps"] = np.ones((n, 2))
        A["r_max"] = np.zeros(n); A["isSphere"] = np.zeros(n, np.int32); A["blockedDOFs"] = np.zeros(n, np.int32)
        A["isDamped"] = np.ones(n, np.int32); A["material"] = np.zeros(n, np.int32)
        A["wallAxis"] = np.zeros(n, np.int32); A["wallSense"] = np.zeros(n, np.int32)
        A["hasBound"] = np.zeros(n, np.int32)
        for i, b in enumerate(self.bodies):
            b.id = i
            s = b.state
            A["pos"][i] = s.pos; A["ori"][i] = s.ori; A["vel"][i] = s.vel; A["angVel"][i] = s.angVel
            A["angMom"][i] = s.angMom; A["mass"][i] = s.mass; A["inertia"][i] = s.inertia
            A["blockedDOFs"][i] = s.blockedDOFs; A["isDamped"][i] = 1 if s.isDamped else 0
            if b.material not in mats:
                raise ValueError(f"body {i}: material not in O.materials")
            A["material"][i] = mats.index(b.material)
            A["hasBound"][i] = 1 if b.hasBound else 0
            if isinstance(b.shape, Wall):
                A["kind"][i] = nat.SDM_BODY_WALL
                A["wallAxis"][i] = b.shape.axis; A["wallSense"][i] = b.shape.sense
            else:
                A["kind"][i] = nat.SDM_BODY_SUPERQUADRIC
                A["rxyz"][i] = b.shape.rxyz; A["eps"][i] = b.shape.eps; A["r_max"][i] = b.shape.r_max
                A["isSphere"][i] = 1 if b.shape.isSphere else 0
        return A

    def _open(self):
        if self.ctx is not None:
            return
        lib = self.lib or nat.load_hip()
        params = engines_to_params(self.engines, self.dt, self.precision,
                                   self.cell.size if self.periodic else None)
        self.ctx = nat.Context(lib, params, len(self.bodies), [m._native() for m in self.materials])
        upload_arrays(self.ctx, self._arrays())
        for b in self.bodies:
            b._scene = self
        for e in self.engines:
            if isinstance(e, CompressionEngine):
                e._attach(self.ctx)

    def step(self):
        self._open()
        self.ctx.step(1)

    def run(self, nSteps=1, wait=True):
        self._open()
        self.ctx.step(int(nSteps))

    @property
    def iter(self):
        return int(self.ctx.get("scene.iter")[0]) if self.ctx else 0

    @property
    def time(self):
        return float(self.ctx.get("scene.time")[0]) if self.ctx else 0.0

    # -- bulk attribute access --
    def get(self, name):
        self._open()
        return self.ctx.get(name)

    def set(self, name, values):
        self._open()
        self.ctx.set(name, values)

    def counters(self):
        return self.ctx.counters()

    def interactions(self):
        """(id1, id2, isReal) of the interaction container in its linear order."""
        self._open()
        i1 = self.ctx.get("intr.id1")
        i2 = self.ctx.get("intr.id2")
        real = (self.ctx.get("intr.hasGeom") != 0) & (self.ctx.get("intr.hasPhys") != 0)
        return i1, i2, real


def open_arrays(lib: nat.SdmLibrary, params: nat.sdm_params, arrays: dict, materials) -> nat.Context:
    """Open a context directly from bulk arrays (used for large synthetic packings)."""
    n = len(arrays["kind"])
    ctx = nat.Context(lib, params, n, [m._native() if hasattr(m, "_native") else m for m in materials])
    upload_arrays(ctx, arrays)
    return ctx
