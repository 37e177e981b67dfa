"""Differentiable rollout over a frozen contact set (the differentiation form), as a
torch.autograd.Function.

``FrozenProblem.from_context`` freezes the contacts of an engine context (their normals,
contact points, depths and shear forces, the bodies' positions, masses and world-frame inertia)
at the current step; ``rollout(problem, params, v0)`` then returns the loss of a K-step window
of the linear law + NewtonIntegrator on that contact set, differentiable with respect to
``params = (kn, ks, mu)`` and the start velocities ``v0``.  Forward and adjoint are HIP kernels
of libsdm_hip.so (``sdm_frozen_forward`` / ``sdm_frozen_backward``); the loss is the mean,
over the last ``loss_steps`` steps, of the summed normal-force magnitudes of the contacts of
``loss_body`` (the moving top wall in the reference workload).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native as nat

_P = ctypes.c_void_p


class sdm_frozen_problem(ctypes.Structure):
    _fields_ = [("n_bodies", ctypes.c_int32), ("n_contacts", ctypes.c_int32),
                ("pos0", _P), ("mass", _P), ("inv_inertia", _P), ("dynamic", _P), ("wall_vel", _P),
                ("id_a", _P), ("id_b", _P), ("normal", _P), ("ra", _P), ("rb", _P), ("depth0", _P), ("fs0", _P),
                ("in_loss", _P), ("csr_off", _P), ("csr_ent", _P),
                ("dt", ctypes.c_double), ("damping", ctypes.c_double), ("gravity", ctypes.c_double * 3),
                ("steps", ctypes.c_int32), ("loss_steps", ctypes.c_int32), ("checkpoint_every", ctypes.c_int32)]


def _bind(lib):
    L = lib.lib
    if getattr(L, "_frozen_bound", False):
        return L
    ptr = ctypes.POINTER(sdm_frozen_problem)
    L.sdm_frozen_checkpoint_size.restype = ctypes.c_int64
    L.sdm_frozen_checkpoint_size.argtypes = [ptr]
    L.sdm_frozen_work_size.restype = ctypes.c_int64
    L.sdm_frozen_work_size.argtypes = [ptr]
    L.sdm_frozen_forward.restype = ctypes.c_int
    L.sdm_frozen_forward.argtypes = [ptr, _P, _P, _P, _P, _P, ctypes.POINTER(ctypes.c_double)]
    L.sdm_frozen_backward.restype = ctypes.c_int
    L.sdm_frozen_backward.argtypes = [ptr, _P, _P, _P, ctypes.POINTER(ctypes.c_double), _P]
    L._frozen_bound = True
    return L


class FrozenProblem:
    """Host arrays of a frozen window plus their device copies (float64)."""

    def __init__(self, arrays: dict, dt: float, damping: float, gravity, steps: int, loss_steps: int,
                 checkpoint_every: int = 50, device="cuda"):
        self.host = arrays
        self.n = len(arrays["mass"])
        self.m = len(arrays["depth0"])
        self.dt, self.damping, self.gravity = float(dt), float(damping), tuple(float(g) for g in gravity)
        self.steps, self.loss_steps, self.checkpoint_every = int(steps), int(loss_steps), int(checkpoint_every)
        self.dev = {}
        for k, v in arrays.items():
            dt_ = torch.int32 if v.dtype.kind in "iu" else torch.float64
            self.dev[k] = torch.as_tensor(np.ascontiguousarray(v), dtype=dt_, device=device).contiguous()
        self.w0 = self.dev["w0"]
        s = sdm_frozen_problem()
        s.n_bodies, s.n_contacts = self.n, self.m
        for k in ("pos0", "mass", "inv_inertia", "dynamic", "wall_vel", "id_a", "id_b", "normal", "ra", "rb",
                  "depth0", "fs0", "in_loss", "csr_off", "csr_ent"):
            setattr(s, k, self.dev[k].data_ptr() if self.dev[k].numel() else None)
        s.dt, s.damping = self.dt, self.damping
        s.gravity[:] = self.gravity
        s.steps, s.loss_steps, s.checkpoint_every = self.steps, self.loss_steps, self.checkpoint_every
        self.struct = s

    @staticmethod
    def arrays_from_context(ctx: nat.Context, loss_body: int, wall_velocity: dict):
        """The frozen window of an engine context at its current state: every interaction
        with geometry, physics and positive depth is a contact (normal, contact point, depth and
        shear force as the last step left them); walls (blockedDOFs = all) are non-dynamic and
        move with ``wall_velocity[id]`` (default 0)."""
        pos = ctx.get("state.pos").reshape(-1, 3)
        n = len(pos)
        mass = ctx.get("state.mass")
        inertia = ctx.get("state.inertia").reshape(-1, 3)
        R = ctx.get("shape.rot_mat2global").reshape(-1, 3, 3)
        blocked = ctx.get("state.blockedDOFs")
        dynamic = (blocked != 63).astype(np.int32)
        with np.errstate(divide="ignore"):
            inv = np.where(inertia > 0, 1.0 / np.where(inertia > 0, inertia, 1.0), 0.0)
        invJ = np.einsum("nij,nj,nkj->nik", R, inv, R).reshape(-1, 9)
        invJ[dynamic == 0] = 0.0
        wall_vel = np.zeros((n, 3))
        for k, v in wall_velocity.items():
            wall_vel[k] = v
        i1, i2 = ctx.get("intr.id1"), ctx.get("intr.id2")
        real = (ctx.get("intr.hasGeom") != 0) & (ctx.get("intr.hasPhys") != 0)
        depth = ctx.get("geom.PenetrationDepth")
        sel = np.nonzero(real & (depth > 0))[0]
        a, b = i1[sel].astype(np.int32), i2[sel].astype(np.int32)
        cp = ctx.get("geom.contactPoint").reshape(-1, 3)[sel]
        normal = ctx.get("geom.normal").reshape(-1, 3)[sel]
        fs = ctx.get("phys.shearForce").reshape(-1, 3)[sel]
        m = len(sel)
        # per-body contact lists in contact order
        ent = np.concatenate([np.arange(m) * 2, np.arange(m) * 2 + 1]).astype(np.int64)
        owner = np.concatenate([a, b]).astype(np.int64)
        order = np.lexsort((ent, owner))
        csr_ent = ent[order].astype(np.int32)
        csr_off = np.zeros(n + 1, np.int32)
        np.add.at(csr_off, owner + 1, 1)
        csr_off = np.cumsum(csr_off).astype(np.int32)
        return {
            "pos0": pos.copy(), "mass": mass.copy(), "inv_inertia": invJ, "dynamic": dynamic, "wall_vel": wall_vel,
            "id_a": a, "id_b": b, "normal": normal, "ra": cp - pos[a], "rb": cp - pos[b], "depth0": depth[sel].copy(),
            "fs0": fs, "in_loss": (a == loss_body).astype(np.int32), "csr_off": csr_off, "csr_ent": csr_ent,
            "v0": ctx.get("state.vel").reshape(-1, 3).copy(), "w0": ctx.get("state.angVel").reshape(-1, 3).copy(),
        }


def ctx_dtype(t: torch.Tensor):
    return t.dtype if t.is_floating_point() else torch.float64


class FrozenRollout(torch.autograd.Function):
    """loss = FrozenRollout.apply(params[3], v0[n, 3], problem) with params = (kn, ks, mu).
    params and v0 may be float32 or float64: the kernels work in double, the loss is double and
    each gradient comes back in the dtype of its input."""

    @staticmethod
    def forward(ctx, params, v0, problem):
        params_in, v0_in = params, v0
        lib = nat.load_hip()
        L = _bind(lib)
        sp = ctypes.byref(problem.struct)
        params = params.detach().to(torch.float64).contiguous()
        v0 = v0.detach().to(torch.float64).contiguous()
        ck = torch.empty(int(L.sdm_frozen_checkpoint_size(sp)), dtype=torch.float64, device=v0.device)
        work = torch.empty(int(L.sdm_frozen_work_size(sp)), dtype=torch.float64, device=v0.device)
        loss = ctypes.c_double()
        torch.cuda.synchronize()
        lib.check(L.sdm_frozen_forward(sp, params.data_ptr(), v0.data_ptr(), problem.w0.data_ptr(), ck.data_ptr(),
                                       work.data_ptr(), ctypes.byref(loss)))
        ctx.problem, ctx.ck, ctx.work, ctx.params = problem, ck, work, params
        ctx.dtypes = (ctx_dtype(params_in), ctx_dtype(v0_in))
        return torch.tensor(loss.value, dtype=torch.float64, device=v0.device)

    @staticmethod
    def backward(ctx, grad_loss):
        lib = nat.load_hip()
        L = _bind(lib)
        problem = ctx.problem
        gp = (ctypes.c_double * 3)()
        gv = torch.empty((problem.n, 3), dtype=torch.float64, device=ctx.ck.device)
        torch.cuda.synchronize()
        lib.check(L.sdm_frozen_backward(ctypes.byref(problem.struct), ctx.params.data_ptr(), ctx.ck.data_ptr(),
                                        ctx.work.data_ptr(), gp, gv.data_ptr()))
        g = torch.tensor(list(gp), dtype=torch.float64, device=gv.device)
        # gradients in the dtypes of the inputs (float32 parameters and velocities are taken in
        # double by the kernels)
        return (grad_loss * g).to(ctx.dtypes[0]), (grad_loss * gv).to(ctx.dtypes[1]), None


def rollout(problem: FrozenProblem, params: torch.Tensor, v0: torch.Tensor) -> torch.Tensor:
    return FrozenRollout.apply(params, v0, problem)
