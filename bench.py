"""Benchmark: particle-steps/s of the superellipsoid DEM step at 200k particles on one MI355X.

Workload (the configuration the headline metric is quoted on): 200,000 superellipsoids,
semi-axes U[0.8,1.2] mm, exponents U[0.6,1.4], density 2650 kg/m^3, a dense bed in a
120x120x200 mm box of six walls, Hertz-Mindlin law (Ig2_*_SuperquadricsGeom2 / Nelder-Mead,
Ip2_SuperquadricsMat2_*, SuperquadricsLaw2; E=7e7 Pa, nu=0.3, mu=0.5), gravity, local damping
0.3, dt = 1e-6 s.  One step = ForceResetter + InsertionSortCollider (when its stride test
fires) + InteractionLoop + NewtonIntegrator.

Procedure: generate the seeded lattice (configs.config1), densify it with the grow-and-relax
packing algorithm run by the engine itself (configs.densify: contacts with ~2 um overlaps,
coordination ~3.5), run W warm-up steps in FP64, snapshot the state, then time K steps of the
all-FP64 form and K steps of the mixed FP32/FP64 form, each from that same snapshot.  The CPU
oracle (the restatement of the reference's OpenMP loop) runs a bounded sample from the same
snapshot on the host cores.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# HBM bytes per launch of k_solve_nm<float> on this workload: rocprofv3 FETCH_SIZE and
# WRITE_SIZE (median over the timed steps), scaled by the same counters on a 1 GiB copy
# (FETCH_SIZE x 2.000, WRITE_SIZE x 1.000); profiles/c1_200k_hbm_traffic.md.
SOLVER_TRAFFIC_BYTES = (221226 * 2.0 + 166754 * 1.0) * 1024
FLOP_EQ_PER_EVAL = 392  # pow=8, sin/cos=4, atan2=16, div=4, add/mul=1 over one Disfun evaluation
PEAK_FP32 = 157.3e12
PEAK_FP64 = 78.6e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


ERROR_FIELDS = ("geom.PenetrationDepth", "geom.normal", "geom.contactPoint", "phys.normalForce", "contact.force",
                "diag.stop")


def mixed_error(ref, mix):
    """Per-contact error of the mixed form against the FP64 path after one step from the same
    state: relative depth error and relative vector errors (|a - b| / |b|) over the contacts
    both forms report, with the share of contacts whose solve took the same path."""
    d64, dmx = ref["geom.PenetrationDepth"], mix["geom.PenetrationDepth"]
    both = (d64 > 0) & (dmx > 0)
    out = {"contacts": int(both.sum()), "classified_differently": int(((d64 > 0) != (dmx > 0)).sum()),
           "same_solver_path": float((ref["diag.stop"][both] == mix["diag.stop"][both]).mean())}

    def stats(err):
        q = np.percentile(err, [50, 99, 99.9]) if len(err) else [0, 0, 0]
        return {"median": float(q[0]), "p99": float(q[1]), "p999": float(q[2]), "max": float(err.max()) if len(err) else 0,
                "frac_le_1e-5": float((err <= 1e-5).mean()) if len(err) else 1.0}
    out["depth"] = stats(np.abs(dmx[both] - d64[both]) / d64[both])
    for k, name in (("geom.normal", "normal"), ("geom.contactPoint", "contact_point"), ("phys.normalForce", "normal_force"),
                    ("contact.force", "force")):
        a = mix[k].reshape(-1, 3)[both]
        b = ref[k].reshape(-1, 3)[both]
        out[name] = stats(np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-300))
    return out


def timed_run(ctx, steps, timers=True):
    """Wall time of `steps` steps.  With timers the phases and the bandwidth-bound kernels are
    bracketed by HIP events on the context's stream (and the step graph is off); without,
    steady-state steps replay the captured graph."""
    ctx.reset_counters()
    ctx.set_timing(timers)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(steps)
    ctx.sync()
    wall = time.perf_counter() - t0
    if not timers:
        return wall, ctx.counters(), None
    tm = ctx.timers()
    tm["kernels"] = ctx.kernel_timers()
    return wall, ctx.counters(), tm


def calibrate_copy(nbytes=1 << 30, reps=4):
    """A plain device-to-device copy of `nbytes`, `reps` times: per dispatch it reads and
    writes exactly nbytes, which scales the HBM byte counters of a profiler run."""
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "copy", "bytes_read_per_dispatch": nbytes, "bytes_written_per_dispatch": nbytes,
                      "dispatches": reps}))


def adjoint_bench(args, n=20000, steps=500, loss_steps=100, checkpoint_every=50, wall_speed=0.05):
    """Differentiation workload: n superellipsoids (the config-0 size and shape distributions)
    densified into a bed, the top wall lowered onto the tallest particles and moving down at
    wall_speed; a `steps`-step frozen-contact rollout whose loss is the mean top-wall normal
    force over the last `loss_steps` steps; gradients w.r.t. (kn, ks, mu) and the 3n start
    velocities, checkpoints every `checkpoint_every` steps.  The kn gradient is checked against
    a central difference of the device forward."""
    import torch

    from sudodem_amd import configs, load_hip
    from sudodem_amd.adjoint import FrozenProblem, rollout
    from sudodem_amd.scene import rot_matrices
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    lib = load_hip()
    t0 = time.perf_counter()
    cfg = configs.densify(configs.config1(n_particles=n, box_xy=40e-3), lib, iterations=100)
    A = cfg.arrays
    part = A["kind"] == 0
    ez = np.zeros((int(part.sum()), 3))
    ez[:, 2] = 1.0
    tops = A["pos"][part, 2] + configs.support_h(A["rxyz"][part], A["eps"][part], rot_matrices(A["ori"][part]), ez)
    A["pos"][5, 2] = float(np.sort(tops)[-40]) - 1e-6  # the 40 tallest particles touch the top wall
    ctx = cfg.open(lib)
    ctx.step(20)
    arr = FrozenProblem.arrays_from_context(ctx, loss_body=5, wall_velocity={5: (0.0, 0.0, -wall_speed)})
    ctx.close()
    P = FrozenProblem(arr, cfg.dt, 0.3, (0.0, 0.0, -9.81), steps, loss_steps, checkpoint_every)
    log("window: %d bodies, %d contacts (%d on the top wall), prepared in %.1fs" % (
        P.n, P.m, int(arr["in_loss"].sum()), time.perf_counter() - t0))
    params0 = (1e5, 7e4, 0.5)
    p = torch.tensor(params0, dtype=torch.float64, device="cuda", requires_grad=True)
    v = torch.tensor(arr["v0"], dtype=torch.float64, device="cuda", requires_grad=True)
    rollout(P, p.detach(), v.detach())  # first call: allocation and code load
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss = rollout(P, p, v)
    torch.cuda.synchronize()
    tf = time.perf_counter() - t0
    t0 = time.perf_counter()
    loss.backward()
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    h = 1e-6 * params0[0]
    lp = float(rollout(P, torch.tensor((params0[0] + h,) + params0[1:], dtype=torch.float64, device="cuda"), v.detach()))
    lm = float(rollout(P, torch.tensor((params0[0] - h,) + params0[1:], dtype=torch.float64, device="cuda"), v.detach()))
    fd = (lp - lm) / (2 * h)
    g = p.grad.cpu().numpy()
    nb = int(arr["dynamic"].sum())
    print(json.dumps({
        "metric": "frozen-contact rollout, forward + adjoint, particle-steps/s",
        "value": nb * steps / (tf + tb), "unit": "particle-steps/s", "n_gpus": 1, "steps": steps, "warmup": 1,
        "ms_per_step": (tf + tb) / steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "%d superellipsoids, %d-step window, top wall at %.2f m/s, loss over the last %d steps, "
                               "checkpoints every %d steps" % (nb, steps, wall_speed, loss_steps, checkpoint_every),
                   "contacts": P.m},
        "forward_ms": tf * 1e3, "backward_ms": tb * 1e3, "loss": float(loss.detach()),
        "grad_kn_ks_mu": g.tolist(), "grad_v0_norm": float(v.grad.norm()),
        "grad_kn_central_difference": fd, "grad_kn_rel_err": abs(g[0] - fd) / max(abs(fd), 1e-300),
    }))


def batch_bench(args):
    """Batching workload: specimens with log-uniform particle counts in [2000, 50000], each with
    its own Kn and friction (linear law, LM geometry), advanced together in one context, against
    the same specimens each in its own context stepped one after the other."""
    import torch

    from sudodem_amd import configs, load_hip
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    lib = load_hip()
    t0 = time.perf_counter()
    specs = configs.config4(n_specimens=args.batch)
    b = configs.batch(specs)
    log("%d specimens, %d bodies, generated in %.1fs" % (len(specs), b.n, time.perf_counter() - t0))
    nparts = sum(c.n - 6 for c in specs)
    K, W = args.steps, args.warmup
    bc = b.open(lib)
    bc.step(W)
    bc.sync()
    from sudodem_amd.autotune import autotune
    t0 = time.perf_counter()
    tuning, trials = autotune(bc, steps=20, log=log)
    log("autotuned in %.1fs: %s" % (time.perf_counter() - t0, tuning))
    bc.reset_counters()
    t0 = time.perf_counter()
    bc.step(K)
    bc.sync()
    tb = time.perf_counter() - t0
    graph_steps = bc.counters()["graph_steps"]
    log("batched: %.3f ms/step (%d of %d steps replayed from the step graph)" % (tb / K * 1e3, graph_steps, K))
    bc.close()
    ctxs = [c.open(lib) for c in specs]
    for c in ctxs:
        c.step(W)
    for c in ctxs:
        c.sync()
    t0 = time.perf_counter()
    for _ in range(K):
        for c in ctxs:
            c.step(1)
    for c in ctxs:
        c.sync()
    ts = time.perf_counter() - t0
    log("one by one: %.3f ms/step" % (ts / K * 1e3))
    print(json.dumps({
        "metric": "aggregate particle-steps/s, %d independent specimens in one context" % len(specs),
        "value": nparts * K / tb, "unit": "particle-steps/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": tb / K * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "%d box specimens (log-uniform 2000-50000 superellipsoids, own Kn and friction, "
                               "linear law)" % len(specs), "particles": nparts, "parallelism": "single GPU"},
        "one_by_one": {"value": nparts * K / ts, "ms_per_step": ts / K * 1e3},
        "batched_over_one_by_one": ts / tb,
        "tuning": tuning, "graph_steps": graph_steps,
        "autotune_trials": [dict(t, ms_per_step=ms * 1e3) for t, ms in trials],
    }))


def _sample_pair_set_check(ctx, cfg, real_before, n_sample=2000, seed=0):
    """Exactness of the periodic pair set on a sample of bodies, right after a collider run:
    for each sampled body, the brute-force set of partners whose boxes overlap it in the
    periodic cell (closed intervals; candidates from a periodic k-d tree in the max norm),
    against the interactions of that body in the container.  Returns the number of missing
    pairs and of extra pairs that were not real contacts when the collider ran."""
    from scipy.spatial import cKDTree
    L = np.array(cfg.cell)
    bmin, bmax = ctx.get("bound.min").reshape(-1, 3), ctx.get("bound.max").reshape(-1, 3)
    c = 0.5 * (bmin + bmax)
    w = bmax - bmin
    wmax = float(w.max())
    tree = cKDTree(np.mod(c, L), boxsize=L)
    rng = np.random.default_rng(seed)
    sample = np.sort(rng.choice(len(c), size=min(n_sample, len(c)), replace=False))
    i1, i2 = ctx.get("intr.id1"), ctx.get("intr.id2")
    insample = np.zeros(len(c), bool)
    insample[sample] = True
    have, have_real = {}, {}
    for a, b in zip(i1.tolist(), i2.tolist()):
        r = (a, b) in real_before or (b, a) in real_before
        for x, y in ((a, b), (b, a)):
            if insample[x]:
                have.setdefault(x, set()).add(y)
                if r:
                    have_real.setdefault(x, set()).add(y)
    missing = extra = checked = 0
    for i in sample:
        cand = tree.query_ball_point(np.mod(c[i], L), r=0.5 * (w[i].max() + wmax) * (1 + 1e-9), p=np.inf)
        cand = np.array([j for j in cand if j != i], dtype=np.int64)
        d = c[cand] - c[i]
        d -= L * np.round(d / L)
        ok = np.all(np.abs(d) <= 0.5 * (w[cand] + w[i]), axis=1)
        expect = set(int(j) for j in cand[ok])
        got = have.get(int(i), set())
        missing += len(expect - got)
        extra += len((got - expect) - have_real.get(int(i), set()))
        checked += len(expect)
    return {"bodies_sampled": int(len(sample)), "overlapping_pairs_checked": checked, "missing": missing,
            "extra_non_real": extra}


def periodic_bench(args):
    """Periodic workload: args.periodic superellipsoids (r lognormal mean 1 mm / log-sigma 0.25,
    rz = U[0.5,1.5] r, eps U[0.3,1.9]) in a cubic periodic cell, linear law (Levenberg-Marquardt
    Geom with the periodic shift), FP64, no gravity.  Seeded lattice + grow-and-relax packing,
    W warm-up steps, K timed steps; pair-set exactness checked on a sample of bodies after the
    timed run; the CPU oracle runs a bounded sample on a 20k-particle cell of the same
    distributions (its initial sweep over a million bodies alone would take hours).  The dense
    packing is an 8000-particle cell compressed under dynamics and tiled (configs.config2_dense)."""
    import torch

    from sudodem_amd import SDM_FP64, configs, load_hip
    from sudodem_amd._native import Context
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    lib = load_hip()
    t0 = time.perf_counter()
    k = max(1, int(round((args.periodic / 8000.0) ** (1.0 / 3.0))))
    cfg = configs.config2_dense(lib, n_particles=args.periodic, base=args.periodic // k ** 3, log=log)
    log("dense periodic packing of %d bodies in %.1fs (cell %.1f mm, solids fraction %.3f)" % (
        cfg.n, time.perf_counter() - t0, cfg.cell[0] * 1e3, cfg.solids_fraction))
    rng = np.random.default_rng(7)
    ctx = cfg.open(lib, precision=SDM_FP64)
    ctx.set("state.vel", rng.normal(scale=0.05, size=(cfg.n, 3)))  # agitation: the collider keeps running
    t0 = time.perf_counter()
    ctx.step(args.warmup)
    ctx.sync()
    log("warm-up %d steps in %.1fs" % (args.warmup, time.perf_counter() - t0))
    K = args.steps
    wall, cnt, tm = timed_run(ctx, K)
    N = cfg.n
    depth = ctx.get("geom.PenetrationDepth")
    # pair-set exactness right after the next collider run the steps trigger themselves
    runs = ctx.counters()["collider_runs"]
    for _ in range(100):
        real = (ctx.get("intr.hasGeom") != 0) & (ctx.get("intr.hasPhys") != 0)
        real_before = set(zip(ctx.get("intr.id1")[real].tolist(), ctx.get("intr.id2")[real].tolist()))
        ctx.step(1)
        if ctx.counters()["collider_runs"] > runs:
            break
    else:
        raise RuntimeError("no collider run within 100 steps")
    check = _sample_pair_set_check(ctx, cfg, real_before=real_before)
    log("pair-set check: %s" % check)
    ctx.close()
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_lib import load_oracle
        orc = load_oracle()
        threads = max(1, min(os.cpu_count() or 1, 16))
        small = cfg.base
        cctx = small.open(orc)
        orc.check(orc.set_threads(cctx.ptr, threads), cctx.ptr)
        cctx.set("state.vel", np.random.default_rng(7).normal(scale=0.05, size=(small.n, 3)))
        cctx.step(1)  # initial sort and sweep outside the sample
        t0 = time.perf_counter()
        csteps = 0
        while time.perf_counter() - t0 < 10.0 and csteps < 5000:  # ~10 s of CPU work
            cctx.step(10)
            csteps += 10
        cw = time.perf_counter() - t0
        cctx.close()
        cpu = {"value": small.n * csteps / cw, "unit": "particle-steps/s", "cores": threads, "kind": "port",
               "sample": "%d steps of the %d-particle base cell of the same packing, same agitation, FP64 "
                         "oracle (restatement of the reference loop), OpenMP %d threads, %.1fs wall" % (
                             csteps, small.n, threads, cw)}
        log("cpu baseline: %.0f particle-steps/s" % cpu["value"])
    print(json.dumps({
        "metric": "particle-steps/s and contact-evals/s, periodic cell, %d superellipsoids" % N,
        "value": N * K / wall, "unit": "particle-steps/s", "n_gpus": 1, "steps": K, "warmup": args.warmup,
        "ms_per_step": wall / K * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "%d superellipsoids (eps U[0.3,1.9], rz/rx U[0.5,1.5], lognormal sizes) in a "
                               "%.1f mm periodic cell (a %d-particle compressed cell tiled %dx%dx%d), linear law "
                               "(LM Geom), dt=1e-6 s" % (N, cfg.cell[0] * 1e3, cfg.base.n, k, k, k),
                   "n_particles": N, "parallelism": "single GPU"},
        "contact_evals_per_s": cnt["contact_evals"] / wall, "dist_evals_per_s": cnt["dist_evals"] / wall,
        "collider_runs": cnt["collider_runs"], "interactions": int(len(depth)), "contacts": int((depth > 0).sum()),
        "solids_fraction": cfg.solids_fraction,
        "phases_ms_per_step": {k: tm[k] / K for k in ("collider_ms", "interactions_ms", "newton_ms", "solver_ms")},
        "pair_set_check": check,
        "cpu_baseline": cpu,
    }))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--n", type=int, default=200000)
    ap.add_argument("--cpu-steps", type=int, default=10, help="steps of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--densify", type=int, default=150, help="iterations of the packing algorithm")
    ap.add_argument("--batch", type=int, default=0, metavar="SPECIMENS",
                    help="batching workload instead: this many box specimens in one context vs one by one")
    ap.add_argument("--adjoint", action="store_true",
                    help="differentiation workload instead: 20k-body frozen-contact rollout, forward + adjoint")
    ap.add_argument("--periodic", type=int, default=0, metavar="PARTICLES",
                    help="periodic workload instead: this many superellipsoids in a periodic cell (1000000)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a device copy of known size (the reference for rocprof FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.batch:
        batch_bench(args)
        return
    if args.adjoint:
        adjoint_bench(args)
        return
    if args.periodic:
        periodic_bench(args)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design note only",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    from sudodem_amd import SDM_FP64, SDM_MIXED, configs, load_hip
    from sudodem_amd._native import Context

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    lib = load_hip()
    t0 = time.perf_counter()
    lattice = configs.config1(n_particles=args.n)
    log("lattice of %d bodies in %.1fs (solids fraction %.3f)" % (lattice.n, time.perf_counter() - t0,
                                                                   lattice.solids_fraction))
    t0 = time.perf_counter()
    cfg = configs.densify(lattice, lib, iterations=args.densify, log=log)
    log("densified in %.1fs (solids fraction %.3f, mean scale %.3f)" % (time.perf_counter() - t0,
                                                                        cfg.solids_fraction, cfg.mean_scale))
    base = cfg.open(lib, precision=SDM_FP64)
    t0 = time.perf_counter()
    base.step(args.warmup)
    base.sync()
    log("warm-up %d steps in %.1fs" % (args.warmup, time.perf_counter() - t0))
    snap = base.snapshot()
    base.close()
    log("contacts after warm-up: %d of %d interactions" % (int((snap["geom.PenetrationDepth"] > 0).sum()),
                                                            len(snap["geom.PenetrationDepth"])))
    depth = snap["geom.PenetrationDepth"]
    n_contacts = int((depth > 0).sum())
    n_pairs = len(depth)

    results = {}
    one_step = {}
    for name, prec in (("fp64", SDM_FP64), ("mixed", SDM_MIXED)):
        ctx = Context(lib, cfg.params(prec), cfg.n, [m._native() for m in cfg.materials])
        ctx.restore(snap)
        ctx.step(2)  # first launches (queue allocation) outside the timed region
        ctx.restore(snap)
        wall, cnt, _ = timed_run(ctx, args.steps, timers=False)
        # the same steps again with per-phase event timers (launched kernels, no graph)
        ctx.restore(snap)
        wall_t, _, tm = timed_run(ctx, args.steps, timers=True)
        results[name] = (wall, cnt, tm)
        log("%s: %.3f ms/step (%d of %d steps replayed from the step graph; %.3f ms/step launched with timers), "
            "solver %.3f ms/step, %d collider runs" % (name, wall / args.steps * 1e3, cnt["graph_steps"], args.steps,
                                                       wall_t / args.steps * 1e3, tm["solver_ms"] / args.steps,
                                                       cnt["collider_runs"]))
        log("%s kernels: %s" % (name, json.dumps(tm["kernels"])))
        log("%s timers: %s" % (name, json.dumps({k: v for k, v in tm.items() if k != "kernels"})))
        # one step from the snapshot again, kept for the per-contact error of the mixed form
        ctx.restore(snap)
        ctx.step(1)
        one_step[name] = {k: ctx.get(k) for k in ERROR_FIELDS}
        ctx.close()

    N = cfg.n - 6
    K = args.steps
    w64, c64, t64 = results["fp64"]
    wmx, cmx, tmx = results["mixed"]
    pss64 = N * K / w64
    pssmx = N * K / wmx
    # dominant kernel: the Nelder-Mead solver of the mixed form
    solver_s = tmx["solver_ms"] * 1e-3
    achieved = tmx["solver_evals"] * FLOP_EQ_PER_EVAL / solver_s if solver_s > 0 else 0.0
    roof = {"bound": "valu", "kernel": "k_solve_nm<float>", "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12,
            "unit": "TFLOP/s", "frac": achieved / PEAK_FP32, "traffic": SOLVER_TRAFFIC_BYTES if args.n == 200000 else None,
            "flop_eq_per_eval": FLOP_EQ_PER_EVAL, "evals_per_step": tmx["solver_evals"] / K,
            "kernel_ms_per_step": tmx["solver_ms"] / K}
    solver64_s = t64["solver_ms"] * 1e-3
    ach64 = t64["solver_evals"] * FLOP_EQ_PER_EVAL / solver64_s if solver64_s > 0 else 0.0

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_lib import load_oracle
        orc = load_oracle()
        threads = max(1, min(os.cpu_count() or 1, 16))
        cctx = Context(orc, cfg.params(SDM_FP64), cfg.n, [m._native() for m in cfg.materials])
        cctx.restore(snap)
        orc.check(orc.set_threads(cctx.ptr, threads), cctx.ptr)
        t0 = time.perf_counter()
        cctx.step(args.cpu_steps)
        cw = time.perf_counter() - t0
        cctx.close()
        cpu = {"value": N * args.cpu_steps / cw, "unit": "particle-steps/s", "cores": threads, "kind": "port",
               "sample": "%d steps of the same 200k-particle snapshot, FP64 oracle (restatement of the "
                         "reference loop), OpenMP %d threads, %.1fs wall" % (args.cpu_steps, threads, cw)}
        log("cpu baseline: %.0f particle-steps/s" % cpu["value"])

    out = {
        "metric": "particle-steps/s at 200k superellipsoids: FP32/FP64 mix vs all-FP64 GPU",
        "value": pssmx,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": wmx / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "200000 superellipsoids, Hertz-Mindlin (Nelder-Mead Geom2 + Law2), dense bed "
                               "(grow-and-relax packing) in a 120x120x200 mm box, dt=1e-6 s", "n_particles": N, "precision": "mixed",
                   "parallelism": "single GPU"},
        "fp64": {"value": pss64, "ms_per_step": w64 / K * 1e3, "solver_ms_per_step": t64["solver_ms"] / K,
                 "solver_achieved_tflops_eq": ach64 / 1e12, "solver_frac_of_fp64_peak": ach64 / PEAK_FP64},
        "mixed_over_fp64": pssmx / pss64,
        "graph_steps": {"fp64": c64["graph_steps"], "mixed": cmx["graph_steps"]},
        "contact_evals_per_s": cmx["contact_evals"] / wmx,
        "evaluated_pairs_per_s": tmx["queued"] / K / (wmx / K),
        "dist_evals_per_s": cmx["dist_evals"] / wmx,
        "interactions": n_pairs,
        "contacts": n_contacts,
        "solids_fraction": cfg.solids_fraction,
        "queued_per_step": {"fp64": t64["queued"] / K, "mixed": tmx["queued"] / K},
        "phases_ms_per_step": {k: tmx[k] / K for k in ("collider_ms", "interactions_ms", "newton_ms", "solver_ms")},
        "kernels_ms_per_step": {name: {k: v["ms"] / K for k, v in tm["kernels"].items() if isinstance(v, dict)}
                                for name, tm in (("fp64", t64), ("mixed", tmx))},
        "mixed_error_vs_fp64": mixed_error(one_step["fp64"], one_step["mixed"]),
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
