"""FP64 CPU forward-mode (tangent) restatement of the frozen-contact rollout, the checker of
the HIP adjoint (sdm_frozen_forward / sdm_frozen_backward).  Test infrastructure only.

Each function propagates a value and one tangent (the directional derivative along a chosen
direction of (kn, ks, mu, v0)); the loss tangent is then dLoss . direction, which the adjoint's
gradient must reproduce.  Same model as the device code: per contact with depth > 0,
Fn = kn depth n, Fs = cap(Fs - ks dt vt) with |Fs| <= mu |Fn|, F = -Fn - Fs applied to body a
(with r_a x F) and -F to body b (with -r_b x F); Newton's leapfrog with Cundall factors held at
their values; angular velocity with the frozen world-frame inverse inertia.
"""
import numpy as np


def _sgn(a):
    return np.sign(a)


def rollout_jvp(A, params, v0, steps, loss_steps, dt, damping, gravity, dparams=(0.0, 0.0, 0.0), dv0=None,
                stats=None):
    """Returns (loss, dloss) along the direction (dparams, dv0).  `stats` (a dict), when given,
    receives the per-contact normal and shear forces of the first step ("fn1", "fs1"), the
    number of contact-steps on the Coulomb cap ("capped") and of Cundall factors other than 1
    ("damped")."""
    kn, ks, mu = params
    dkn, dks, dmu = dparams
    n, m = len(A["mass"]), len(A["depth0"])
    dyn = A["dynamic"].astype(bool)
    x = A["pos0"].copy()
    v = np.where(dyn[:, None], v0, A["wall_vel"])
    w = np.where(dyn[:, None], A["w0"], 0.0)
    fs = A["fs0"].copy()
    dx = np.zeros((n, 3))
    dv = np.where(dyn[:, None], dv0 if dv0 is not None else 0.0, 0.0)
    dw = np.zeros((n, 3))
    dfs = np.zeros((m, 3))
    a_, b_ = A["id_a"].astype(int), A["id_b"].astype(int)
    nrm, ra, rb = A["normal"], A["ra"], A["rb"]
    invJ = A["inv_inertia"].reshape(-1, 3, 3)
    g = np.asarray(gravity, float)
    loss = dloss = 0.0
    inl = A["in_loss"].astype(bool)
    ent = A["csr_ent"].astype(np.int64)
    ce, side = ent >> 1, ent & 1
    owner = np.repeat(np.arange(n), np.diff(A["csr_off"]))
    arm = np.where(side[:, None] == 0, ra[ce], rb[ce])
    for t in range(steps):
        ua, ub = x[a_] - A["pos0"][a_], x[b_] - A["pos0"][b_]
        dua, dub = dx[a_], dx[b_]
        depth = A["depth0"] - np.einsum("ij,ij->i", nrm, ub - ua)
        ddepth = -np.einsum("ij,ij->i", nrm, dub - dua)
        act = depth > 0
        vrel = (v[b_] + np.cross(w[b_], rb)) - (v[a_] + np.cross(w[a_], ra))
        dvrel = (dv[b_] + np.cross(dw[b_], rb)) - (dv[a_] + np.cross(dw[a_], ra))
        vt = vrel - np.einsum("ij,ij->i", nrm, vrel)[:, None] * nrm
        dvt = dvrel - np.einsum("ij,ij->i", nrm, dvrel)[:, None] * nrm
        fnm = kn * depth
        dfnm = dkn * depth + kn * ddepth
        ftr = fs - ks * dt * vt
        dftr = dfs - dks * dt * vt - ks * dt * dvt
        ft = np.linalg.norm(ftr, axis=1)
        cap = act & (ft > mu * fnm) & (ft > 0)
        fsn = ftr.copy()
        dfsn = dftr.copy()
        if cap.any():
            s = ftr[cap] / ft[cap][:, None]
            ds_dir = dftr[cap]
            dft = np.einsum("ij,ij->i", s, ds_dir)
            rho = mu * fnm[cap] / ft[cap]
            drho = (dmu * fnm[cap] + mu * dfnm[cap]) / ft[cap] - mu * fnm[cap] * dft / ft[cap] ** 2
            fsn[cap] = rho[:, None] * ftr[cap]
            dfsn[cap] = drho[:, None] * ftr[cap] + rho[:, None] * dftr[cap]
        if stats is not None:
            if t == 0:
                stats["fn1"] = np.where(act[:, None], fnm[:, None] * nrm, 0.0)
                stats["fs1"] = np.where(act[:, None], fsn, 0.0)
            stats["capped"] = stats.get("capped", 0) + int(cap.sum())
        F = -(fnm[:, None] * nrm) - fsn
        dF = -(dfnm[:, None] * nrm) - dfsn
        F[~act] = 0
        dF[~act] = 0
        fsn[~act] = 0
        dfsn[~act] = 0
        if t >= steps - loss_steps:
            loss += fnm[act & inl].sum() / loss_steps
            dloss += dfnm[act & inl].sum() / loss_steps
        # forces and torques per body, in contact order per body: np.add.at adds the entries in
        # array order, and the entries of one body are contiguous in list order
        Fb = np.zeros((n, 3))
        dFb = np.zeros((n, 3))
        Tb = np.zeros((n, 3))
        dTb = np.zeros((n, 3))
        sgn_e = np.where(side == 0, 1.0, -1.0)[:, None]
        np.add.at(Fb, owner, sgn_e * F[ce])
        np.add.at(dFb, owner, sgn_e * dF[ce])
        np.add.at(Tb, owner, sgn_e * np.cross(arm, F[ce]))
        np.add.at(dTb, owner, sgn_e * np.cross(arm, dF[ce]))
        m_ = A["mass"][:, None]
        acc = Fb / np.where(m_ > 0, m_, 1.0) + g
        dacc = dFb / np.where(m_ > 0, m_, 1.0)
        ca = 1 - damping * _sgn(acc * (v + 0.5 * dt * acc))
        cT = 1 - damping * _sgn(Tb * w)
        if stats is not None:
            stats["damped"] = stats.get("damped", 0) + int(((ca != 1) & dyn[:, None]).sum() + ((cT != 1) & dyn[:, None]).sum())
        vn = np.where(dyn[:, None], v + dt * ca * acc, v)
        dvn = np.where(dyn[:, None], dv + dt * ca * dacc, 0.0)
        wn = np.where(dyn[:, None], w + dt * np.einsum("nij,nj->ni", invJ, cT * Tb), w)
        dwn = np.where(dyn[:, None], dw + dt * np.einsum("nij,nj->ni", invJ, cT * dTb), 0.0)
        x = x + dt * vn
        dx = dx + dt * dvn
        v, dv, w, dw, fs, dfs = vn, dvn, wn, dwn, fsn, dfsn
    return loss, dloss
