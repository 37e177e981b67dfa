"""Benchmark: particle-steps/s of the superellipsoid DEM step at 200k particles on one MI355X.

Workload (the configuration the headline metric is quoted on): 200,000 superellipsoids,
semi-axes U[0.8,1.2] mm, exponents U[0.6,1.4], density 2650 kg/m^3, a dense bed in a
120x120x200 mm box of six walls, Hertz-Mindlin law (Ig2_*_SuperquadricsGeom2 / Nelder-Mead,
Ip2_SuperquadricsMat2_*, SuperquadricsLaw2; E=7e7 Pa, nu=0.3, mu=0.5), gravity, local damping
0.3, dt = 1e-6 s.  One step = ForceResetter + InsertionSortCollider (when its stride test
fires) + InteractionLoop + NewtonIntegrator.

Procedure: generate the seeded lattice (configs.config1), densify it with the grow-and-relax
packing algorithm run by the engine itself (configs.densify: contacts with ~2 um overlaps,
coordination ~3.3, solids fraction 0.573), run W warm-up steps in FP64, snapshot the state, then time K steps of the
all-FP64 form and K steps of the mixed FP32/FP64 form, each from that same snapshot.  The CPU
oracle (the restatement of the reference's OpenMP loop) runs a bounded sample from the same
snapshot on the host cores (median of 5 short repeats).  Prints one JSON line, with the
device time of every kernel of the step (HIP events on the step's stream) and roofline
fractions of the dominant VALU-bound kernel and of the HBM-bound integrator, bound update
and collider.
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
SOLVER_TRAFFIC_BYTES = (229236 * 2.0 + 99257 * 1.0) * 1024
FLOP_EQ_PER_EVAL = 392  # pow=8, sin/cos=4, atan2=16, div=4, add/mul=1 over one Disfun evaluation
PEAK_FP32 = 157.3e12
PEAK_FP64 = 78.6e12
PEAK_HBM = 8.0e12
# compulsory HBM bytes: NewtonIntegrator reads F, T, v, L, q, x, m, I and flags (188 B) and writes
# v, L, omega, q, x (128 B) per body, and refreshes the 112 B of the packed body record the
# narrowphase reads (position, velocities, rotation); the bound update reads x and r_max and
# writes min / max (80 B) and the record's bound line (48 B); the collider reads the bounds
# (48 B per body) and writes 8 B per pair
NEWTON_BYTES_PER_BODY = 316 + 112
BOUNDS_BYTES_PER_BODY = 80 + 48


def host_cpu():
    """CPU model name of the host (first 'model name' of /proc/cpuinfo) and its logical CPUs."""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return name, os.cpu_count()


def cpu_threads():
    """Threads of the CPU baseline: OMP_NUM_THREADS when set (the host share given to this
    process), else every CPU this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


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
    """Differentiation workload: n superellipsoids (the plain workload's size and shape
    distributions, linear law Kn 1e5 / Ks 7e4 / mu 0.5) settled in a box by the wall servo
    (configs.config3: 20 kPa on every wall, 20,000 steps), then the top wall driven down at
    wall_speed: a `steps`-step frozen-contact rollout (local damping 0.3, gravity) whose loss is
    the mean top-wall normal force over the last `loss_steps` steps; gradients w.r.t. (kn, ks,
    mu) and the 3n start velocities, all float32 at the interface (the kernels work in double),
    checkpoints every `checkpoint_every` steps.  The three parameter gradients and a random
    directional derivative of the velocity gradient are checked against the FP64 forward-mode
    restatement (tests/frozen_ref.py: the exact derivative of the same discrete rollout) and
    against central differences of the device forward (the best of several step sizes: the loss
    is piecewise smooth, every contact leaving the window and every Cundall sign switch is a
    jump)."""
    import torch

    from sudodem_amd import configs, load_hip
    from sudodem_amd.adjoint import FrozenProblem, rollout
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    lib = load_hip()
    t0 = time.perf_counter()
    ctx, cfg = configs.config3(lib, n_particles=n, box_xy=40e-3, wall_speed=wall_speed, log=log)
    cn = ctx.mean_coordination()
    arr = FrozenProblem.arrays_from_context(ctx, loss_body=configs.TOP_WALL,
                                            wall_velocity={configs.TOP_WALL: (0.0, 0.0, -wall_speed)})
    ctx.close()
    P = FrozenProblem(arr, cfg.dt, 0.3, (0.0, 0.0, -9.81), steps, loss_steps, checkpoint_every)
    log("window: %d bodies, %d contacts (%d on the top wall), coordination %.2f, prepared in %.1fs" % (
        P.n, P.m, int(arr["in_loss"].sum()), cn, time.perf_counter() - t0))
    params0 = (1e5, 7e4, 0.5)
    p = torch.tensor(params0, dtype=torch.float32, device="cuda", requires_grad=True)
    v = torch.tensor(arr["v0"], dtype=torch.float32, device="cuda", requires_grad=True)
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
    g = p.grad.double().cpu().numpy()
    gv = v.grad.double().cpu().numpy()
    pd = p.detach().double()
    vd = v.detach().double()
    v32 = vd.cpu().numpy()  # the float32 start velocities, as the kernels took them
    dyn = arr["dynamic"] == 1
    rng = np.random.default_rng(11)
    dv = np.zeros_like(arr["v0"])
    dv[dyn] = rng.normal(size=(int(dyn.sum()), 3))
    dad = float(np.sum(gv * dv))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from frozen_ref import rollout_jvp
    # exact derivatives of the same discrete rollout (Coulomb cap and Cundall branches held at
    # their values) by the FP64 forward-mode restatement on the host, one direction per pass
    t0 = time.perf_counter()
    jvp = []
    for k in range(3):
        d = [0.0, 0.0, 0.0]
        d[k] = 1.0
        jvp.append(rollout_jvp(arr, params0, v32, steps, loss_steps, cfg.dt, 0.3, (0.0, 0.0, -9.81), dparams=d)[1])
    jv = rollout_jvp(arr, params0, v32, steps, loss_steps, cfg.dt, 0.3, (0.0, 0.0, -9.81), dv0=dv)[1]
    tjvp = time.perf_counter() - t0
    fm_rel = [abs(g[k] - jvp[k]) / max(abs(jvp[k]), 1e-300) for k in range(3)] + [abs(dad - jv) / max(abs(jv), 1e-300)]

    def L(pp, vv):
        return float(rollout(P, pp, vv))

    def fd_best(f, x0, tangent, rel_steps):
        # the loss is piecewise smooth (contacts leaving the window, Cundall sign switches): the
        # smallest difference over the step sizes, of the steps that straddle no such event
        errs = []
        for r in rel_steps:
            h = r * x0
            errs.append(abs((f(h) - f(-h)) / (2 * h) - tangent) / max(abs(tangent), 1e-300))
        return min(errs)
    fd_rel = []
    for k in range(3):
        def fk(h, k=k):
            e = torch.zeros(3, dtype=torch.float64, device="cuda")
            e[k] = h
            return L(pd + e, vd)
        fd_rel.append(fd_best(fk, params0[k], g[k], (1e-3, 1e-4, 1e-5, 1e-6)))
    dvt = torch.tensor(dv, dtype=torch.float64, device="cuda")
    fd_rel.append(fd_best(lambda h: L(pd, vd + h * dvt), 1.0, dad, (1e-4, 1e-5, 1e-6, 1e-7)))
    # the same window without local damping (no Cundall switches): gradients against central
    # differences where the loss is smooth but for contacts leaving the window
    P0 = FrozenProblem(arr, cfg.dt, 0.0, (0.0, 0.0, -9.81), steps, loss_steps, checkpoint_every)
    p0 = pd.clone().requires_grad_(True)
    v0t = vd.clone().requires_grad_(True)
    l0 = rollout(P0, p0, v0t)
    l0.backward()
    g0, gv0 = p0.grad.cpu().numpy(), v0t.grad.cpu().numpy()
    fd0_rel = []
    for k in range(3):
        def fk0(h, k=k):
            e = torch.zeros(3, dtype=torch.float64, device="cuda")
            e[k] = h
            return float(rollout(P0, pd + e, vd))
        fd0_rel.append(fd_best(fk0, params0[k], g0[k], (1e-3, 1e-4, 1e-5, 1e-6)))
    fd0_rel.append(fd_best(lambda h: float(rollout(P0, pd, vd + h * dvt)), 1.0, float(np.sum(gv0 * dv)),
                           (1e-4, 1e-5, 1e-6, 1e-7)))
    log("loss %.6g; gradients (kn, ks, mu) %s, velocity direction %.6g; forward-mode rel. differences %s (%.1fs); "
        "central differences, best step: %s; without damping: %s" % (float(loss.detach()), g.tolist(), dad, fm_rel,
                                                                      tjvp, fd_rel, fd0_rel))
    cpu = None
    if not args.no_cpu:
        cs = 20
        t0 = time.perf_counter()
        rollout_jvp(arr, params0, v32, cs, 1, cfg.dt, 0.3, (0.0, 0.0, -9.81), dparams=(1.0, 0.0, 0.0))
        cw = time.perf_counter() - t0
        nb0 = int(dyn.sum())
        cpu = {"value": nb0 * cs / cw, "unit": "particle-steps/s", "cores": 1, "kind": "port",
               "sample": "%d steps of the same window, forward + one tangent direction, FP64 forward-mode "
                         "restatement of the frozen rollout (numpy, one thread), %.1fs wall; a full velocity "
                         "gradient by forward mode would take 3n such passes" % (cs, cw)}
    nb = int(dyn.sum())
    print(json.dumps({
        "metric": "frozen-contact rollout, forward + adjoint, particle-steps/s",
        "value": nb * steps / (tf + tb), "unit": "particle-steps/s", "n_gpus": 1, "steps": steps, "warmup": 1,
        "ms_per_step": (tf + tb) / steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (f32 parameters and velocities at the interface)", "data": "synthetic",
        "config": {"workload": "%d superellipsoids settled by the wall servo (20 kPa), %d-step window, top wall at "
                               "%.2f m/s, local damping 0.3, loss over the last %d steps, checkpoints every %d steps"
                               % (nb, steps, wall_speed, loss_steps, checkpoint_every),
                   "contacts": P.m, "top_wall_contacts": int(arr["in_loss"].sum()), "coordination": cn},
        "forward_ms": tf * 1e3, "backward_ms": tb * 1e3, "loss": float(loss.detach()),
        "grad_kn_ks_mu": g.tolist(), "grad_v0_norm": float(np.linalg.norm(gv)), "grad_v0_direction": dad,
        "forward_mode_rel_err": {"kn": fm_rel[0], "ks": fm_rel[1], "mu": fm_rel[2], "v0_direction": fm_rel[3]},
        "central_difference_rel_err_best_step": {"kn": fd_rel[0], "ks": fd_rel[1], "mu": fd_rel[2],
                                                 "v0_direction": fd_rel[3]},
        "central_difference_rel_err_best_step_no_damping": {"kn": fd0_rel[0], "ks": fd0_rel[1], "mu": fd0_rel[2],
                                                            "v0_direction": fd0_rel[3], "loss": float(l0)},
        "cpu_baseline": cpu,
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
    # the same number of steps again with kernel timers: the dominant kernel's roofline
    _, _, tm = timed_run(bc, K, timers=True)
    bc.close()
    # Specimens at rest: the certified screen of the first evaluation leaves the LM solver a few
    # hundred evaluations per step, and the step is the interaction pass and the integrator;
    # the integrator's HBM roofline (compulsory bytes per body) is the one with a fixed count
    km = tm["kernels"]
    nb = nparts + 6 * len(specs)
    nms = km["k_newton"]["ms"] / max(1, km["k_newton"]["launches"])
    ach = NEWTON_BYTES_PER_BODY * nb / (nms * 1e-3) if nms > 0 else 0.0
    roof = {"bound": "hbm", "kernel": "k_newton", "achieved": ach / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
            "frac": ach / PEAK_HBM, "traffic": None, "algorithmic_bytes_per_launch": NEWTON_BYTES_PER_BODY * nb,
            "bytes": "%d B per body (316 B state + 112 B record)" % NEWTON_BYTES_PER_BODY, "ms_per_launch": nms,
            "solver": {"kernel": "k_solve_lm (FP64 Levenberg-Marquardt, main launch)",
                       "evals_per_step": tm["solver_evals"] / K, "kernel_ms_per_step": tm["solver_ms"] / K}}
    kernels = {k: v["ms"] / K for k, v in tm["kernels"].items() if isinstance(v, dict)}
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
    for c in ctxs:
        c.close()
    cpu = None
    if not args.no_cpu:
        # bounded sample: the first specimens, each alone on the oracle, from their warm state
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_lib import load_oracle
        orc = load_oracle()
        threads = cpu_threads()
        model, ncpu = host_cpu()
        done_ps, cw, used = 0, 0.0, 0
        for c in specs:
            cctx = c.open(orc)
            orc.check(orc.set_threads(cctx.ptr, threads), cctx.ptr)
            cctx.step(2)  # initial sort and sweep outside the sample
            t0 = time.perf_counter()
            cctx.step(20)
            cw += time.perf_counter() - t0
            cctx.close()
            done_ps += (c.n - 6) * 20
            used += 1
            if cw > 10.0:
                break
        cpu = {"value": done_ps / cw, "unit": "particle-steps/s", "cores": threads, "kind": "port",
               "host": {"model": model, "logical_cpus": ncpu},
               "sample": "20 steps of each of the first %d specimens alone (%d particles), FP64 oracle (restatement "
                         "of the reference loop), %d threads, %.1fs wall" % (used, done_ps // 20, threads, cw)}
        log("cpu baseline: %.0f particle-steps/s" % cpu["value"])
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
        "kernels_ms_per_step": kernels,
        "roofline": roof,
        "cpu_baseline": cpu,
    }))


def _sample_pair_set_check(ctx, cfg, real_before, n_sample=2000, seed=0):
    """The periodic pair set on a sample of bodies, right after a collider run: for each sampled
    body, the brute-force set of partners whose boxes overlap it in the periodic cell (closed
    intervals; candidates from a periodic k-d tree in the max norm), against the interactions
    of that body in the container.  Returns the number of missing pairs (the reference's walk
    from vector index 0 can leave a crossing for the next run: the caller checks that run) and
    of extra pairs that were not real contacts when the collider ran, and the missing pairs."""
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
    missing_pairs = []
    for i in sample:
        cand = tree.query_ball_point(np.mod(c[i], L), r=0.5 * (w[i].max() + wmax) * (1 + 1e-9), p=np.inf)
        cand = np.array([j for j in cand if j != i], dtype=np.int64)
        d = c[cand] - c[i]
        d -= L * np.round(d / L)
        ok = np.all(np.abs(d) <= 0.5 * (w[cand] + w[i]), axis=1)
        expect = set(int(j) for j in cand[ok])
        got = have.get(int(i), set())
        missing += len(expect - got)
        missing_pairs += [(int(i), j) for j in sorted(expect - got)]
        extra += len((got - expect) - have_real.get(int(i), set()))
        checked += len(expect)
    return {"bodies_sampled": int(len(sample)), "overlapping_pairs_checked": checked, "missing": missing,
            "extra_non_real": extra}, missing_pairs


def _late_pairs_resolved(ctx, cfg, pairs):
    """After the next collider run: how many of `pairs` still overlap and are still absent."""
    L = np.array(cfg.cell)
    bmin, bmax = ctx.get("bound.min").reshape(-1, 3), ctx.get("bound.max").reshape(-1, 3)
    have = {(min(a, b), max(a, b)) for a, b in zip(ctx.get("intr.id1").tolist(), ctx.get("intr.id2").tolist())}
    bad = 0
    for a, b in pairs:
        c = 0.5 * (bmin[[a, b]] + bmax[[a, b]])
        w = bmax[[a, b]] - bmin[[a, b]]
        d = c[1] - c[0]
        d -= L * np.round(d / L)
        bad += bool(np.all(np.abs(d) <= 0.5 * (w[0] + w[1]))) and (min(a, b), max(a, b)) not in have
    return bad


def periodic_bench(args):
    """Periodic workload: args.periodic superellipsoids (r lognormal mean 1 mm / log-sigma 0.25,
    rz = U[0.5,1.5] r, eps U[0.3,1.9]) in a cubic periodic cell, linear law (Levenberg-Marquardt
    Geom with the periodic shift), FP64 (--mixed: the mixed form), no gravity.  Seeded lattice +
    grow-and-relax packing,
    W warm-up steps, K timed steps; pair-set exactness checked on a sample of bodies after the
    timed run; the CPU oracle runs a bounded sample on a 20k-particle cell of the same
    distributions (its initial sweep over a million bodies alone would take hours).  The dense
    packing is an 8000-particle cell compressed under dynamics and tiled (configs.config2_dense)."""
    import torch

    from sudodem_amd import SDM_FP64, SDM_MIXED, configs, load_hip
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
    ctx = cfg.open(lib, precision=SDM_MIXED if args.mixed else SDM_FP64)
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
    check, late = _sample_pair_set_check(ctx, cfg, real_before=real_before)
    runs = ctx.counters()["collider_runs"]
    for _ in range(100):
        ctx.step(1)
        if ctx.counters()["collider_runs"] > runs:
            break
    check["late_pairs_missing_after_next_run"] = _late_pairs_resolved(ctx, cfg, late)
    log("pair-set check: %s" % check)
    ctx.close()
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_lib import load_oracle
        orc = load_oracle()
        threads = cpu_threads()
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
        "dtype": "f32+f64" if args.mixed else "f64", "data": "synthetic",
        "config": {"workload": "%d superellipsoids (eps U[0.3,1.9], rz/rx U[0.5,1.5], lognormal sizes) in a "
                               "%.1f mm periodic cell (a %d-particle compressed cell tiled %dx%dx%d), linear law "
                               "(LM Geom), dt=1e-6 s" % (N, cfg.cell[0] * 1e3, cfg.base.n, k, k, k),
                   "n_particles": N, "parallelism": "single GPU"},
        "contact_evals_per_s": cnt["contact_evals"] / wall, "dist_evals_per_s": cnt["dist_evals"] / wall,
        "collider_runs": cnt["collider_runs"], "interactions": int(len(depth)), "contacts": int((depth > 0).sum()),
        "solids_fraction": cfg.solids_fraction,
        "phases_ms_per_step": {k: tm[k] / K for k in ("collider_ms", "interactions_ms", "newton_ms", "solver_ms")},
        "kernels_ms_per_step": {k: v["ms"] / K for k, v in tm["kernels"].items() if isinstance(v, dict)},
        "roofline": {"bound": "valu", "kernel": "k_solve_lm (%s Levenberg-Marquardt, main launch)" % (
                         "mixed" if args.mixed else "FP64"),
                     "achieved": tm["solver_evals"] * FLOP_EQ_PER_EVAL / (tm["solver_ms"] * 1e-3) / 1e12,
                     "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                     "frac": tm["solver_evals"] * FLOP_EQ_PER_EVAL / (tm["solver_ms"] * 1e-3) / PEAK_FP64,
                     "traffic": None, "flop_eq_per_eval": FLOP_EQ_PER_EVAL, "evals_per_step": tm["solver_evals"] / K,
                     "kernel_ms_per_step": tm["solver_ms"] / K},
        "pair_set_check": check,
        "cpu_baseline": cpu,
    }))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--n", type=int, default=200000)
    ap.add_argument("--cpu-steps", type=int, default=4, help="steps of each CPU baseline repeat")
    ap.add_argument("--cpu-repeats", type=int, default=5, help="repeats of the CPU baseline sample (median)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--densify", type=int, default=150, help="iterations of the packing algorithm")
    ap.add_argument("--batch", type=int, default=0, metavar="SPECIMENS",
                    help="batching workload instead: this many box specimens in one context vs one by one")
    ap.add_argument("--adjoint", action="store_true",
                    help="differentiation workload instead: 20k-body frozen-contact rollout, forward + adjoint")
    ap.add_argument("--mixed", action="store_true", help="with --periodic: the mixed precision form")
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
            "traffic_source": "rocprofv3 FETCH_SIZE / WRITE_SIZE of this kernel on this workload, scaled by a copy "
                              "of known size (profiles/c1_200k_hbm_traffic.md); not measured by this run",
            "flop_eq_per_eval": FLOP_EQ_PER_EVAL, "evals_per_step": tmx["solver_evals"] / K,
            "kernel_ms_per_step": tmx["solver_ms"] / K}
    km = tmx["kernels"]
    nb = cfg.n
    runs = max(1, cmx["collider_runs"])
    def hbm(name, bytes_per, ms, launches, note):
        ach = bytes_per / (ms * 1e-3 / max(1, launches)) if ms > 0 else 0.0
        return {"kernel": name, "bound": "hbm", "algorithmic_bytes_per_launch": bytes_per,
                "ms_per_launch": ms / max(1, launches), "achieved": ach / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                "frac": ach / PEAK_HBM, "bytes": note}
    rooflines = [
        hbm("k_newton", NEWTON_BYTES_PER_BODY * nb, km["k_newton"]["ms"], km["k_newton"]["launches"],
            "%d B per body (316 B state + 112 B record)" % NEWTON_BYTES_PER_BODY),
        hbm("k_bounds", BOUNDS_BYTES_PER_BODY * nb, km["k_bounds"]["ms"], km["k_bounds"]["launches"],
            "%d B per body per collider run" % BOUNDS_BYTES_PER_BODY),
        hbm("collider run", 48 * nb + 8 * n_pairs, tmx["collider_ms"], runs,
            "48 B per body + 8 B per pair per run; time = device events around the whole run, host "
            "synchronisations included"),
    ]
    solver64_s = t64["solver_ms"] * 1e-3
    ach64 = t64["solver_evals"] * FLOP_EQ_PER_EVAL / solver64_s if solver64_s > 0 else 0.0

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_lib import load_oracle
        orc = load_oracle()
        threads = cpu_threads()
        model, ncpu = host_cpu()
        cctx = Context(orc, cfg.params(SDM_FP64), cfg.n, [m._native() for m in cfg.materials])
        orc.check(orc.set_threads(cctx.ptr, threads), cctx.ptr)
        walls = []
        for _ in range(args.cpu_repeats):  # each repeat from the same snapshot
            cctx.restore(snap)
            t0 = time.perf_counter()
            cctx.step(args.cpu_steps)
            walls.append(time.perf_counter() - t0)
        cctx.close()
        cw = float(np.median(walls))
        cpu = {"value": N * args.cpu_steps / cw, "unit": "particle-steps/s", "cores": threads, "kind": "port",
               "host": {"model": model, "logical_cpus": ncpu},
               "sample": "median of %d repeats of %d steps from the same 200k-particle snapshot (%s s wall each), "
                         "FP64 oracle (restatement of the reference loop, OpenMP per-thread force buffers as the "
                         "reference's), %d threads" % (args.cpu_repeats, args.cpu_steps,
                                                       ", ".join("%.2f" % w for w in walls), threads)}
        log("cpu baseline: %.0f particle-steps/s (%s, %d threads)" % (cpu["value"], model, threads))

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
        "evals_note": "contact_evals_per_s counts the Ig2 invocations the reference makes (every interaction "
                      "with a bounding-sphere overlap or a stored geometry, including those the step skips "
                      "because their outcome is provably unchanged); evaluated_pairs_per_s counts the pairs "
                      "whose first vertex the step actually evaluates; dist_evals_per_s the distance evaluations",
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
        "rooflines_hbm": rooflines,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
