"""ctypes binding of the C ABI in include/sdm.h.

`SdmLibrary(path, prefix)` binds one shared library that exports the ABI under the given
symbol prefix.  The product library is libsdm_hip.so (prefix ``sdm_``), built in-tree next to
this file; `load_hip()` loads it and raises if it is missing, so nothing silently falls back to
another implementation.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
HIP_LIB = os.path.join(PKG_DIR, "libsdm_hip.so")

SDM_LAW_LINEAR, SDM_LAW_HERTZ = 0, 1
SDM_FP64, SDM_MIXED = 0, 1
SDM_BODY_SUPERQUADRIC, SDM_BODY_WALL = 0, 1
SDM_F64, SDM_I32, SDM_I64 = 0, 1, 2
_NP_DTYPE = {SDM_F64: np.float64, SDM_I32: np.int32, SDM_I64: np.int64}


class sdm_params(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_double),
        ("gravity", ctypes.c_double * 3),
        ("damping", ctypes.c_double),
        ("exact_aspherical_rot", ctypes.c_int32),
        ("verlet_dist", ctypes.c_double),
        ("target_interv", ctypes.c_int32),
        ("min_sweep_dist_factor", ctypes.c_double),
        ("sort_axis", ctypes.c_int32),
        ("law", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("periodic", ctypes.c_int32),
        ("cell", ctypes.c_double * 3),
    ]


class sdm_material(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("density", "Kn", "Ks", "frictionAngle", "betan", "betas", "G", "nu")]


class sdm_counters(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in
                ("steps", "collider_runs", "contact_evals", "dist_evals", "pairs_inserted", "pairs_erased",
                 "graph_steps")]


class sdm_tuning(ctypes.Structure):
    _fields_ = [("solver_blocks_per_cu", ctypes.c_int32), ("sort_by_cost", ctypes.c_int32),
                ("misc_blocks_per_cu", ctypes.c_int32), ("graph", ctypes.c_int32)]


class sdm_servo(ctypes.Structure):
    _fields_ = [("goal", ctypes.c_double * 3), ("max_vel", ctypes.c_double), ("gain_alpha", ctypes.c_double),
                ("f_threshold", ctypes.c_double), ("fmin_threshold", ctypes.c_double), ("unbf_tol", ctypes.c_double),
                ("iterate_num_max", ctypes.c_int32), ("solve_num_max", ctypes.c_int32),
                ("activated", ctypes.c_int32 * 4), ("z_servo", ctypes.c_int32), ("hydro_strain", ctypes.c_int32),
                ("hydro_strain_rate", ctypes.c_double), ("shear_mode", ctypes.c_int32), ("ramp_interval", ctypes.c_int32),
                ("ramp_chunks", ctypes.c_int32), ("echo_interval", ctypes.c_int32), ("target_strain", ctypes.c_double),
                ("keep_state", ctypes.c_int32), ("continue_flag", ctypes.c_int32), ("size0", ctypes.c_double * 3)]


class sdm_servo_state(ctypes.Structure):
    _fields_ = [("wall_stress", ctypes.c_double * 6), ("stress", ctypes.c_double * 3), ("gain", ctypes.c_double * 6),
                ("stiffness", ctypes.c_double * 6), ("size", ctypes.c_double * 3), ("mean_stress", ctypes.c_double),
                ("deviator_stress", ctypes.c_double), ("unbalanced_force", ctypes.c_double),
                ("iterate_num", ctypes.c_int32), ("solve_num", ctypes.c_int32), ("completed_iter", ctypes.c_int64),
                ("size0", ctypes.c_double * 3), ("ramp_num", ctypes.c_int32), ("pad_", ctypes.c_int32)]


# CompressionEngine attribute defaults (pkg/dem/Compression.hpp:160-210) for the fields of sdm_servo
SERVO_DEFAULTS = dict(goal=(0.0, 0.0, 0.0), max_vel=1.0, gain_alpha=0.5, f_threshold=0.01, fmin_threshold=0.01,
                      unbf_tol=0.01, iterate_num_max=100, solve_num_max=2000, activated=(1, 1, 1, 1), z_servo=1,
                      hydro_strain=0, hydro_strain_rate=0.0, shear_mode=1, ramp_interval=10000, ramp_chunks=400,
                      echo_interval=1000, target_strain=0.15, keep_state=0, continue_flag=0, size0=(0.0, 0.0, 0.0))


class SdmError(RuntimeError):
    pass


class SdmLibrary:
    """One implementation of the ABI (symbols `<prefix>create`, `<prefix>step`, ...)."""

    def __init__(self, path: str, prefix: str):
        if not os.path.exists(path):
            raise SdmError(f"shared library {path} is missing; build it first (__graft_entry__.build())")
        self.path = path
        self.prefix = prefix
        self.lib = ctypes.CDLL(path)
        L, p = self.lib, prefix
        c_ctx = ctypes.c_void_p

        def fn(name, res, args):
            f = getattr(L, p + name)
            f.restype = res
            f.argtypes = args
            return f

        self.version = fn("version", ctypes.c_char_p, [])
        self.create = fn("create", c_ctx, [ctypes.POINTER(sdm_params), ctypes.c_int32, ctypes.c_int32])
        self.destroy = fn("destroy", None, [c_ctx])
        self.last_error = fn("last_error", ctypes.c_char_p, [c_ctx])
        self.set_materials = fn("set_materials", ctypes.c_int, [c_ctx, ctypes.POINTER(sdm_material), ctypes.c_int32])
        self.superquadric_props = fn("superquadric_props", ctypes.c_int,
                                     [ctypes.c_double] * 5 + [ctypes.POINTER(ctypes.c_double)] * 3)
        self.superquadric_props_n = fn("superquadric_props_n", ctypes.c_int,
                                       [ctypes.c_int64] + [ctypes.POINTER(ctypes.c_double)] * 5)
        self.field_info = fn("field_info", ctypes.c_int,
                             [c_ctx, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64)])
        self.get = fn("get", ctypes.c_int, [c_ctx, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int64])
        self.set = fn("set", ctypes.c_int, [c_ctx, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int64])
        self.field_list = fn("field_list", ctypes.c_char_p, [])
        self.step = fn("step", ctypes.c_int, [c_ctx, ctypes.c_int32])
        self.sync = fn("sync", ctypes.c_int, [c_ctx])
        self.get_counters = fn("get_counters", ctypes.c_int, [c_ctx, ctypes.POINTER(sdm_counters)])
        self.reset_counters = fn("reset_counters", ctypes.c_int, [c_ctx])
        self.set_timing = fn("set_timing", ctypes.c_int, [c_ctx, ctypes.c_int32])
        self.set_graph = fn("set_graph", ctypes.c_int, [c_ctx, ctypes.c_int32])
        d = ctypes.POINTER(ctypes.c_double)
        self.mean_coordination = fn("mean_coordination", ctypes.c_int, [c_ctx, d])
        self.stress = fn("stress", ctypes.c_int, [c_ctx, ctypes.c_double, d])
        self.body_forces = fn("body_forces", ctypes.c_int, [c_ctx, ctypes.POINTER(ctypes.c_int32), ctypes.c_int32, d])
        self.set_tuning = fn("set_tuning", ctypes.c_int, [c_ctx, ctypes.POINTER(sdm_tuning)])
        self.get_tuning = fn("get_tuning", ctypes.c_int, [c_ctx, ctypes.POINTER(sdm_tuning)])
        self.get_timers = fn("get_timers", ctypes.c_int, [c_ctx, ctypes.POINTER(ctypes.c_double)])
        self.get_kernel_timers = fn("get_kernel_timers", ctypes.c_int, [c_ctx, ctypes.POINTER(ctypes.c_double)])
        self.set_servo = fn("set_servo", ctypes.c_int, [c_ctx, ctypes.POINTER(sdm_servo)])
        self.get_servo = fn("get_servo", ctypes.c_int, [c_ctx, ctypes.POINTER(sdm_servo_state)])

    def fields(self):
        return self.field_list().decode().split()

    def check(self, rc, ctx=None):
        if rc != 0:
            raise SdmError((self.last_error(ctx) or b"").decode())

    def props(self, rx, ry, rz, e1, e2):
        r = ctypes.c_double()
        v = ctypes.c_double()
        inert = (ctypes.c_double * 3)()
        rc = self.superquadric_props(rx, ry, rz, e1, e2, ctypes.byref(r), ctypes.byref(v), inert)
        self.check(rc)
        return r.value, v.value, np.array(list(inert))

    def props_n(self, rxyz, eps):
        """r_max, volume and inertia (per unit density) of n bodies in one call."""
        rxyz = np.ascontiguousarray(rxyz, dtype=np.float64).reshape(-1, 3)
        eps = np.ascontiguousarray(eps, dtype=np.float64).reshape(-1, 2)
        n = len(rxyz)
        rmax, vol, inert = np.empty(n), np.empty(n), np.empty((n, 3))
        d = ctypes.POINTER(ctypes.c_double)
        rc = self.superquadric_props_n(n, rxyz.ctypes.data_as(d), eps.ctypes.data_as(d), rmax.ctypes.data_as(d),
                                       vol.ctypes.data_as(d), inert.ctypes.data_as(d))
        self.check(rc)
        return rmax, vol, inert


class Context:
    """An open scene on one implementation of the ABI."""

    def __init__(self, lib: SdmLibrary, params: sdm_params, n_bodies: int, materials):
        self.lib = lib
        self.ptr = lib.create(ctypes.byref(params), int(n_bodies), len(materials))
        if not self.ptr:
            raise SdmError((lib.last_error(None) or b"").decode())
        arr = (sdm_material * max(1, len(materials)))()
        for i, m in enumerate(materials):
            arr[i] = m
        lib.check(lib.set_materials(self.ptr, arr, len(materials)), self.ptr)

    def close(self):
        if self.ptr:
            self.lib.destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self, name):
        dt = ctypes.c_int32()
        n = ctypes.c_int64()
        self.lib.check(self.lib.field_info(self.ptr, name.encode(), ctypes.byref(dt), ctypes.byref(n)), self.ptr)
        return dt.value, n.value

    def get(self, name):
        dt, n = self.info(name)
        out = np.empty(n, dtype=_NP_DTYPE[dt])
        self.lib.check(self.lib.get(self.ptr, name.encode(), out.ctypes.data_as(ctypes.c_void_p), n), self.ptr)
        return out

    def set(self, name, values):
        dt, _ = self.info(name)
        a = np.ascontiguousarray(np.asarray(values, dtype=_NP_DTYPE[dt]).ravel())
        self.lib.check(self.lib.set(self.ptr, name.encode(), a.ctypes.data_as(ctypes.c_void_p), a.size), self.ptr)

    def step(self, n=1):
        self.lib.check(self.lib.step(self.ptr, int(n)), self.ptr)

    def sync(self):
        self.lib.check(self.lib.sync(self.ptr), self.ptr)

    def counters(self):
        c = sdm_counters()
        self.lib.check(self.lib.get_counters(self.ptr, ctypes.byref(c)), self.ptr)
        return {k: getattr(c, k) for k, _ in sdm_counters._fields_}

    def reset_counters(self):
        self.lib.check(self.lib.reset_counters(self.ptr), self.ptr)

    def mean_coordination(self) -> float:
        """Shop__meanCoordinationNumber (getMeanCN) of the current state."""
        v = ctypes.c_double()
        self.lib.check(self.lib.mean_coordination(self.ptr, ctypes.byref(v)), self.ptr)
        return v.value

    def stress(self, volume=0.0) -> np.ndarray:
        """Shop::getStress (Love-Weber stress tensor, 3x3) of the current state."""
        out = np.zeros(9)
        self.lib.check(self.lib.stress(self.ptr, float(volume), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                       self.ptr)
        return out.reshape(3, 3)

    def body_forces(self, ids) -> np.ndarray:
        """Forces on the given bodies after the last step (n x 3)."""
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        out = np.zeros((len(ids), 3))
        self.lib.check(self.lib.body_forces(self.ptr, ids.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(ids),
                                            out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), self.ptr)
        return out

    def set_servo(self, enabled=True, **kw):
        """Wall stress servo (CompressionEngine, cubic specimen: consolidation, hydroConsolidation,
        triaxial shear); keyword arguments are the engine's attributes (goal=(goalx, goaly, goalz),
        max_vel, gain_alpha, z_servo, shear_mode, ...).  keep_state=1 changes the attributes of the
        running engine instead of starting a new one."""
        if not enabled:
            self.lib.check(self.lib.set_servo(self.ptr, None), self.ptr)
            return
        p = {**SERVO_DEFAULTS, **kw}
        unknown = set(p) - set(SERVO_DEFAULTS)
        if unknown:
            raise SdmError("unknown servo attributes: %s" % sorted(unknown))
        s = sdm_servo()
        s.goal = (ctypes.c_double * 3)(*[float(x) for x in p["goal"]])
        s.activated = (ctypes.c_int32 * 4)(*[int(x) for x in p["activated"]])
        s.size0 = (ctypes.c_double * 3)(*[float(x) for x in p["size0"]])
        for k in ("max_vel", "gain_alpha", "f_threshold", "fmin_threshold", "unbf_tol", "hydro_strain_rate",
                  "target_strain"):
            setattr(s, k, float(p[k]))
        for k in ("iterate_num_max", "solve_num_max", "z_servo", "hydro_strain", "shear_mode", "ramp_interval",
                  "ramp_chunks", "echo_interval", "keep_state", "continue_flag"):
            setattr(s, k, int(p[k]))
        self.lib.check(self.lib.set_servo(self.ptr, ctypes.byref(s)), self.ptr)

    def servo_state(self) -> dict:
        st = sdm_servo_state()
        self.lib.check(self.lib.get_servo(self.ptr, ctypes.byref(st)), self.ptr)
        out = {}
        for k, t in sdm_servo_state._fields_:
            v = getattr(st, k)
            out[k] = np.array(list(v)) if hasattr(t, "_length_") else v
        return out

    def tuning(self) -> dict:
        t = sdm_tuning()
        self.lib.check(self.lib.get_tuning(self.ptr, ctypes.byref(t)), self.ptr)
        return {k: int(getattr(t, k)) for k, _ in sdm_tuning._fields_}

    def set_tuning(self, **kw):
        t = sdm_tuning(**{**self.tuning(), **kw})
        self.lib.check(self.lib.set_tuning(self.ptr, ctypes.byref(t)), self.ptr)

    def set_graph(self, on=True):
        """Replay steady-state steps as a captured graph (the default)."""
        self.lib.check(self.lib.set_graph(self.ptr, 1 if on else 0), self.ptr)

    def set_timing(self, on=True):
        self.lib.check(self.lib.set_timing(self.ptr, 1 if on else 0), self.ptr)

    def timers(self):
        """Per-phase device times [ms] and solver evaluation counts since set_timing(True)."""
        out = (ctypes.c_double * 8)()
        self.lib.check(self.lib.get_timers(self.ptr, out), self.ptr)
        return {"collider_ms": out[0], "interactions_ms": out[1], "newton_ms": out[2], "steps": int(out[3]),
                "solver_ms": out[4], "solver_evals": int(out[5]), "extra_evals": int(out[6]),
                "queued": int(out[7])}

    KERNEL_TIMERS = ("k_newton", "k_force_sum", "k_bounds", "k_pairs_prep", "queue", "k_pairs_misc", "k_nm_first",
                     "nm_list", "solver", "resume_restarts", "k_nm_refine", "k_pairs_finish", "ring_walk",
                     "ring_windows")

    def kernel_timers(self):
        """Device time [ms] and launches of each kernel (or group of launches) of the step since
        set_timing(True): the integrator, the force sums, the bound update, the interaction-loop
        dispatch pass, the solver queue, the wall and sphere pairs, the first evaluation (the
        Nelder-Mead first vertex or the LM screen) and the solver list, the main solver launch
        (the LM restarts included), the mixed Nelder-Mead resume launch, the mixed form's double
        refinement and the finish; with the mixed Hertz form, the Nelder-Mead
        pairs listed for the solver and the pairs handed to its resume launch; with either mixed
        form, the solver's evaluations done in double (anchors included); and the launches
        that extended a sequential phase of the periodic collider's ring walk."""
        n = len(self.KERNEL_TIMERS)
        out = (ctypes.c_double * (2 * n + 4))()
        self.lib.check(self.lib.get_kernel_timers(self.ptr, out), self.ptr)
        r = {name: {"ms": out[2 * k], "launches": int(out[2 * k + 1])} for k, name in enumerate(self.KERNEL_TIMERS)}
        r["nm_listed"] = int(out[2 * n])
        r["nm_double_evals"] = int(out[2 * n + 1])
        r["nm_handed_over"] = int(out[2 * n + 2])
        r["ring_extensions"] = int(out[2 * n + 3])
        return r

    def snapshot(self):
        """Complete state (bodies, interactions, collider and integrator flags) as arrays."""
        return {name: self.get(name) for name in self.lib.fields()
                if not (self.info(name)[1] == 0 and name in ("collider.bounds", "collider.periods"))}

    def restore(self, snap):
        """Load a snapshot taken from any implementation of the ABI."""
        self.set("intr.id1", snap["intr.id1"])
        for name, val in snap.items():
            if name in ("intr.id1",) or name.startswith("contact.") or name.startswith("diag."):
                continue
            if name.startswith("forces."):
                continue
            self.set(name, val)


_HIP = None


def load_hip() -> SdmLibrary:
    """The MI355X library.  Raises SdmError when it has not been built."""
    global _HIP
    if _HIP is None:
        _HIP = SdmLibrary(HIP_LIB, "sdm_")
    return _HIP
