"""Throughput of the FP64 USL step on one MI355X.

Workload: the square-silo discharge -- 512 x 512 cells (513^2 nodes),
1,048,576 particles (4 x 4 per cell, seeded jitter), g_local_mu2, square_silo
walls, dt = 2e-6.  A step is one explicit_mpm_step_usl_threaded +
time_varying_loads over all particles, with the state resident in HBM.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]
                    [--no-subrecords] [--workload silo|heap|ensemble|adjoint]

Prints one JSON line.  `value` is particle-steps per second on the GPU over
K steps timed between two HIP events (nothing else recorded in between);
`ms_per_step_wall` is the host clock around the same call, a cross-check.
`roofline` prices the dominant kernel's algorithmic bytes (FP64 SoA, each
field moved once per pass that needs it) against the 8 TB/s HBM peak, with
its duration from the start and end events each kernel's own dispatch
records (hipExtLaunchKernelGGL: no event packet between the kernels) in a
second pass of K steps; their sum per step falls short of the span by the
launch gaps between the kernels.  The CPU-baseline leg first gates accuracy (GPU against the
oracle's parity build after 1 and 100 steps) and then times the CPU
restatement of the reference step (reference AoS layout, pthreads, colour
lists, barriers; compiled with the reference's -O3 and default FP
contraction) on the host's cores over a bounded sample of the same
workload.  The sub-records `batched` and `adjoint` (with
`batched_vs_plain` and `adjoint_vs_plain`) measure the 256-member ensemble
against the same members as single jobs and the 200-step adjoint against
its plain forward run, each with its own accuracy check; `nonlocal` times
100 steps of the g_nonlocal fluidity heap (256x256 cells, 262,144
particles) with its CG iterations and final residual.
"""
import argparse
import json
import os
import re
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(REPO, "mpm-2d_amd"), os.path.join(REPO, "oracle")]

import numpy as np  # noqa: E402

import pympm  # noqa: E402
import workloads  # noqa: E402

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per unit (FP64 SoA, each field once per pass that needs it):
#   P2G+grid kernel: particles read x y vx vy m v sxx sxy syy (72 B);
#                    nodes: scatter writes m px py fx fy (40 B), the grid
#                    update reads them (40 B) and writes px py fx fy (32 B)
#   G2P+stress kernel: particles read x y vx vy m v Txx Txy Tyy Tzz gp ux uy
#                    (104 B) and write x y vx vy v Txx Txy Tyy Tzz gp gdp ux uy
#                    (104 B); nodes read m px py fx fy (40 B)
P2G_BYTES_PER_PARTICLE, P2G_BYTES_PER_NODE = 72, 112
G2P_BYTES_PER_PARTICLE, G2P_BYTES_PER_NODE = 208, 40


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def read_traffic(kernel, summary="hbm_traffic_silo_512.txt"):
    """HBM bytes per launch from a committed rocprofv3 PMC summary, if any."""
    path = os.path.join(REPO, "profiles", summary)
    if not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"\s*%s\s+traffic_bytes_per_launch\s*=\s*([0-9.eE+]+)" % kernel, line)
        if m:
            return float(m.group(1))
    return None


def _traffic_rate(traffic_bytes, kernel_ms):
    """Measured HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE,
    calibrated) over the launch's duration, GB/s; None without a summary."""
    if traffic_bytes is None or not kernel_ms:
        return None
    return traffic_bytes / (kernel_ms * 1e-3) / 1e9


def _time_oracle(w, variant, threads, budget_s, log_path=None):
    """(particle-steps/s, steps, seconds) of the CPU restatement stepping
    for `budget_s` after one warm-up step."""
    import pyoracle
    o = pyoracle.from_workload(w, variant=variant, num_threads=threads,
                               log_crossings=log_path is not None, log_path=log_path)
    o.run(1)   # first-touch of the AoS arrays and colour lists
    steps, t0 = 0, time.perf_counter()
    while True:
        o.run(1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 2000:
            break
    nact = int(o.particles()["active"].sum())
    o.close()
    return nact * steps / el, steps, el


def cpu_baseline(w, budget_s=15.0, variants=True):
    """The CPU restatement on all host cores, until `budget_s` of stepping
    (the reported baseline: the fastest CPU setting, logging off).  With
    `variants`, also one thread, and all threads with the reference's
    per-crossing log lines written to a file (libmpm/process_usl.c:651-663),
    each for a third of the budget."""
    flags = open("/proc/cpuinfo").read() if os.path.exists("/proc/cpuinfo") else ""
    variant = "v4" if " avx512f" in flags else "v3"
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    model = re.search(r"model name\s*:\s*(.*)", flags)
    value, steps, el = _time_oracle(w, variant, threads, budget_s)
    solver = (", nonlocal solve: banded Cholesky on the host (the reference uses "
              "CXSparse LU, absent here)" if w.get("nonlocal_material") else "")
    out = {"value": value, "unit": "particle-steps/s", "cores": threads, "kind": "port",
           "sample": (f"{w['name']}: {len(w['particles']['x'])} particles, {steps} steps "
                      f"after 1 warm-up step ({el:.1f} s), oracle/usl_oracle.c x86-64-"
                      f"{variant} build with the reference's -O3 and FP contraction, "
                      f"{threads} pthreads, crossing log off{solver}"),
           "cpu": model.group(1).strip() if model else "unknown"}
    if variants:
        import tempfile
        v1, s1, e1 = _time_oracle(w, variant, 1, budget_s / 3)
        with tempfile.NamedTemporaryFile(prefix="mpm_crossings_", suffix=".log") as lf:
            vl, sl, el_ = _time_oracle(w, variant, threads, budget_s / 3, log_path=lf.name)
            log_bytes = os.path.getsize(lf.name)
        out["variants"] = [
            {"value": v1, "cores": 1, "logging": False,
             "sample": f"{s1} steps ({e1:.1f} s) on 1 thread"},
            {"value": vl, "cores": threads, "logging": True,
             "sample": f"{sl} steps ({el_:.1f} s) on {threads} threads, {log_bytes} bytes of "
                       "crossing log lines"}]
    return out


def bench_heap(args, emit=True):
    """The nonlocal fluidity heap (g_nonlocal, 256x256 cells, 262,144
    particles): the same step plus the per-step fluidity assembly, a
    preconditioned conjugate-gradient solve to |r| <= 1e-10 |f| and the
    stress update.  The solve works on a system that stays in the last-level
    cache, so it is timed and reported, not priced against HBM."""
    w = workloads.heap()
    npart = len(w["particles"]["x"])
    nnodes = w["N"] * w["N"]
    log(f"bench: {w['name']} particles={npart} nodes={nnodes} warmup={args.warmup} "
        f"steps={args.steps}")
    cfg = pympm.make_config(w["N"], w["h"], w["dt"], material=w["material"],
                            bc_kind=w["bc_kind"], bc_fp64=w["bc_fp64"], bc_int=w["bc_int"],
                            g_mag=w["g_mag"], ramp_time=w["ramp_time"],
                            nonlocal_material=w["nonlocal_material"], cg_rel_tol=1e-10)
    sim = pympm.Sim(cfg, w["particles"])
    sim.step(args.warmup)
    sim.sync()
    iters = []
    t0 = time.perf_counter()
    ms = np.zeros(3)
    for _ in range(args.steps):
        ms += sim.time_steps(1)
        iters.append(sim.solver_stats()[0])
    wall = time.perf_counter() - t0
    nact = int(sim.active().sum())
    res = sim.solver_stats()[1]
    result = {
        "metric": "particle-steps/sec FP64 USL step, % HBM roofline; batched+adjoint vs plain",
        "value": nact * args.steps / (ms[0] * 1e-3),
        "unit": "particle-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms[0] / args.steps,
        "ms_per_step_wall": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded triangular pile, lithostatic initial stress)",
        "config": {"workload": "nonlocal fluidity heap, 256x256 cells (257^2 nodes), "
                               f"{npart} particles, g_nonlocal A=0.48 mu_s=0.38, dt=5e-6, "
                               "CG rel tol 1e-10",
                   "particles": npart, "nodes": nnodes, "parallelism": "single GPU"},
        "kernels_ms": {"k_p2g_grid": ms[1] / args.steps,
                       "g2p+assembly+cg+stress": ms[2] / args.steps},
        "cg_iterations_mean": float(np.mean(iters)), "cg_iterations_max": int(np.max(iters)),
        "cg_final_rel_residual": res,
        "roofline": None,
        "cpu_baseline": None,
    }
    sim.close()
    if not args.no_cpu_baseline:
        log("bench: CPU baseline (oracle restatement with a banded Cholesky solve)")
        result["cpu_baseline"] = cpu_baseline(w, args.cpu_budget)
        result["speedup_vs_cpu"] = result["value"] / result["cpu_baseline"]["value"]
    if emit:
        print(json.dumps(result), flush=True)
    return result


def bench_ensemble(args, emit=True):
    """256 independent column collapses of uneven size (33^2 to 181^2
    nodes, 2,048 to 64,800 particles, 4,661,824 in all; per-member mu_s and
    E) stepped together by shared launches replayed from a HIP graph, after
    the bin-capacity autotuner has picked its capacity.  Beside it, the same
    members as 256 single-job contexts (two launches per member per step,
    each context on its own stream)."""
    ens = workloads.ensemble()
    counts = [len(w["particles"]["x"]) for w in ens]
    nodes = [w["N"] * w["N"] for w in ens]
    log(f"bench: ensemble members={len(ens)} particles={sum(counts)} warmup={args.warmup} "
        f"steps={args.steps}")
    cfgs = [pympm.make_config(w["N"], w["h"], w["dt"], material=w["material"],
                              bc_kind=w["bc_kind"], bc_fp64=w["bc_fp64"], bc_int=w["bc_int"],
                              g_mag=w["g_mag"], ramp_time=w["ramp_time"]) for w in ens]
    b = pympm.Batch(cfgs, [w["particles"] for w in ens])
    b.step(args.warmup)
    tile, cap, roles, trial = b.autotune_tiles(trial_steps=4)
    tile_name = {1: "16x8", 2: "8x4", 3: "16x8/512"}[tile]
    log(f"bench: autotuned P2G tiles {tile_name} nodes, bin capacity {cap}, {roles} threads per "
        f"node (ms/step for 16x8 then 8x4, each 16/32/64 x 2/3: {trial.tolist()})")
    nact0 = sum(int(b.active(k).sum()) for k in range(len(ens)))
    t0 = time.perf_counter()
    ms = b.time_steps(args.steps)
    wall = time.perf_counter() - t0
    nact1 = sum(int(b.active(k).sum()) for k in range(len(ens)))
    acc_fields = ("x", "y", "vx", "vy", "txx", "txy", "tyy")
    batched_state = [{f: b.get(k, f) for f in acc_fields} for k in range(len(ens))]
    b.close()
    pstep = (nact0 + nact1) / 2.0 * args.steps
    np_sum, nn_sum = sum(counts), sum(nodes)
    p2g_bytes = P2G_BYTES_PER_PARTICLE * np_sum + P2G_BYTES_PER_NODE * nn_sum
    g2p_bytes = G2P_BYTES_PER_PARTICLE * np_sum + G2P_BYTES_PER_NODE * nn_sum
    kernels = {"k_p2g_grid": (ms[1], p2g_bytes), "k_g2p_stress": (ms[2], g2p_bytes)}
    dom = max(kernels, key=lambda k: kernels[k][0])
    dom_ms, dom_bytes = kernels[dom]
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    step_bytes = p2g_bytes + g2p_bytes
    step_ms = ms[0] / args.steps

    # the same members as single-job contexts, stepped round-robin for as
    # many steps as the batch took, so that the end states must agree
    plain_steps = args.steps
    sims = [pympm.Sim(c, w["particles"]) for c, w in zip(cfgs, ens)]
    for s in sims:
        s.step(args.warmup)
    for s in sims:
        s.sync()
    t1 = time.perf_counter()
    for _ in range(plain_steps):
        for s in sims:
            s.step(1)
    for s in sims:
        s.sync()
    plain_wall = time.perf_counter() - t1
    plain_act = sum(int(s.active().sum()) for s in sims)
    mismatched = [k for k, s in enumerate(sims)
                  if not all(np.array_equal(s.get(f), batched_state[k][f]) for f in acc_fields)]
    for s in sims:
        s.close()
    plain_value = plain_act * plain_steps / plain_wall

    result = {
        "metric": "particle-steps/sec FP64 USL step, % HBM roofline; batched+adjoint vs plain",
        "value": pstep / (ms[0] * 1e-3),
        "unit": "particle-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": step_ms,
        "ms_per_step_wall": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded lattice + jitter per member)",
        "config": {"workload": "ensemble of 256 column collapses, 33^2-181^2 nodes, "
                               f"{np_sum} particles, mu_s~U[0.3,0.5], E~logU[5e5,5e6], "
                               "dt=1e-5",
                   "members": len(ens), "particles": np_sum, "nodes": nn_sum,
                   "p2g_tile_nodes": tile_name, "bin_capacity": cap,
                   "threads_per_node": roles,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": read_traffic(dom, "hbm_traffic_ensemble_256.txt"),
                     "traffic_GBs": _traffic_rate(read_traffic(dom, "hbm_traffic_ensemble_256.txt"),
                                                  dom_ms),
                     "kernel": dom, "kernel_ms": dom_ms, "algorithmic_bytes_per_launch": dom_bytes},
        "kernels_ms": {k: v[0] for k, v in kernels.items()},
        "step_roofline": {"algorithmic_bytes": step_bytes,
                          "achieved_GBs": step_bytes / (step_ms * 1e-3) / 1e9,
                          "frac": step_bytes / (step_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "autotune_ms_per_step": {f"tile{t}_cap{c}_threads{r}": float(trial[6 * i + 2 * k + j])
                                 for i, t in enumerate(("16x8", "8x4"))
                                 for k, c in enumerate((16, 32, 64))
                                 for j, r in enumerate((2, 3))},
        "plain": {"value": plain_value, "steps": plain_steps,
                  "how": "256 single-job contexts, one stream each, stepped round-robin"},
        "batched_vs_plain": (pstep / (ms[0] * 1e-3)) / plain_value,
        "accuracy": {"checked": f"every member after {args.warmup + args.steps} steps, fields "
                                f"{list(acc_fields)}, against the single-job contexts",
                     "members_bitwise_equal": len(ens) - len(mismatched),
                     "members": len(ens), "passed": not mismatched},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        log("bench: CPU baseline (oracle restatement on a sample of members)")
        result["cpu_baseline"] = cpu_baseline_members(ens[::32], args.cpu_budget)
        result["speedup_vs_cpu"] = result["value"] / result["cpu_baseline"]["value"]
    if emit:
        print(json.dumps(result), flush=True)
    return result


def bench_adjoint(args, emit=True):
    """The adjoint workload (129^2 nodes, 65,536 particles, 200 steps,
    checkpoint every 20): a forward run keeping checkpoints, then the reverse
    sweep of L = (front - front*)^2 + (cx - cx*)^2 + (cy - cy*)^2 (runout
    front and particle centroid, calibrate.py) against a run with the target
    constants (mu_s = 0.42, E = 2e6).  A particle-step is one forward step
    and its reverse.  Beside it, the plain forward run (mpm_hip_step, device
    time of the 200 steps).  Accuracy: the reverse sweep's dL/dmu_s, dL/dE
    and dL/d(initial velocities) along a seeded direction against central
    differences of the forward run (relative step 1e-6 for the constants,
    1e-6 m/s along the direction; bound 1e-5)."""
    import calibrate
    w = workloads.adjoint_block()
    steps = w["steps"]
    npart = len(w["particles"]["x"])
    log(f"bench: {w['name']} particles={npart} steps={steps} repeats={args.repeats}")

    def cfg_for(mat):
        return pympm.make_config(w["N"], w["h"], w["dt"], material=mat, bc_kind=w["bc_kind"],
                                 bc_fp64=w["bc_fp64"], bc_int=w["bc_int"], g_mag=w["g_mag"],
                                 ramp_time=w["ramp_time"])
    tsim = pympm.Sim(cfg_for(workloads.ADJOINT_TARGET), w["particles"])
    tsim.step(steps)
    target, _ = calibrate.observables(tsim.get("x"), tsim.get("y"), tsim.active())
    tsim.close()
    warm = pympm.Sim(cfg_for(w["material"]), w["particles"])
    warm.step(steps)
    warm.sync()
    warm.close()
    plain = pympm.Sim(cfg_for(w["material"]), w["particles"])
    t_plain = plain.time_span(steps) * 1e-3     # the 200 steps from the initial state
    plain.close()

    adj = pympm.Adjoint(cfg_for(w["material"]), w["particles"], checkpoint_every=20)
    adj.forward(steps)
    adj.backward({"x": np.zeros(npart)})    # warm-up
    t_f = t_b = 0.0
    for _ in range(args.repeats):
        t0 = time.perf_counter()
        adj.forward(steps)
        t1 = time.perf_counter()
        xf, yf, af = adj.get("x"), adj.get("y"), adj.active()
        loss, seed = calibrate.loss_and_seed(xf, yf, af, target)
        t2 = time.perf_counter()
        grad, gi = adj.backward(seed)
        t3 = time.perf_counter()
        t_f += t1 - t0
        t_b += t3 - t2
    t_f /= args.repeats
    t_b /= args.repeats
    total = t_f + t_b

    # central differences of the forward run: the constants at relative
    # steps 1e-5 (the bound's) and 1e-6 (reported), the initial velocities
    # of the front particle and the largest gradient components at 1e-6 m/s
    def loss_at(mat=None, particles=None):
        if mat is not None:
            adj.set_material(**mat)
        if particles is not None:
            adj.set_particles(particles)
        adj.forward(steps)
        val, _ = calibrate.loss_and_seed(adj.get("x"), adj.get("y"), adj.active(), target)
        return val
    base = dict(w["material"])
    fd = {}
    for k in ("mu_s", "E"):
        for rs in (1e-5, 1e-6):
            hk = rs * abs(base[k])
            lp = loss_at(dict(base, **{k: base[k] + hk}))
            lm = loss_at(dict(base, **{k: base[k] - hk}))
            fd[f"{k} (h={rs:g}|{k}|)"] = ((lp - lm) / (2 * hk), grad[k], rs == 1e-5)
    adj.set_material(**base)
    p0 = w["particles"]
    front = calibrate.observables(xf, yf, af)[1]
    picks = [("vx", front)] \
        + [("vx", int(i)) for i in np.argsort(-np.abs(gi["vx"]))[:2]] \
        + [("vy", int(i)) for i in np.argsort(-np.abs(gi["vy"]))[:2]]
    for f, i in picks:
        vals = []
        for sgn in (1, -1):
            p2 = {q: v.copy() for q, v in p0.items()}
            p2[f][i] += sgn * 1e-6
            vals.append(loss_at(particles=p2))
        fd[f"{f}0[{i}] (h=1e-6)"] = ((vals[0] - vals[1]) / 2e-6, float(gi[f][i]), True)
    adj.set_particles(p0)
    rel = {k: float(abs(a - b) / max(abs(a), 1e-300)) for k, (a, b, _) in fd.items()}
    # the calibration the gradient is for: Gauss-Newton from (0.38, 1e6)
    t0 = time.perf_counter()
    cal = calibrate.calibrate(adj, steps, target, base["mu_s"], base["E"])
    cal_s = time.perf_counter() - t0
    adj.close()

    result = {
        "metric": "particle-steps/sec FP64 USL step, % HBM roofline; batched+adjoint vs plain",
        "value": npart * steps / total,
        "unit": "particle-steps/s (forward + reverse)",
        "n_gpus": 1, "steps": steps, "warmup": 1,
        "ms_per_step": total * 1e3 / steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded lattice + jitter, lithostatic stress, initial slide)",
        "config": {"workload": "adjoint block, 128x128 cells (129^2 nodes), 65,536 particles, "
                               "200 steps, checkpoint every 20, g_local_mu2 "
                               "(mu_s 0.38 vs target 0.42, E 1e6 vs target 2e6), loss = "
                               "squared error of the final runout front and centroid",
                   "particles": npart, "nodes": w["N"] ** 2, "parallelism": "single GPU"},
        "forward_ms": t_f * 1e3, "reverse_ms": t_b * 1e3, "plain_forward_ms": t_plain * 1e3,
        "adjoint_vs_plain": t_plain / total,
        "loss": loss, "dL_dmu_s": grad["mu_s"], "dL_dE": grad["E"],
        "accuracy": {"central_differences": {k: {"fd": float(a), "adjoint": float(b),
                                                 "rel_err": rel[k],
                                                 "gated": g}
                                             for k, (a, b, g) in fd.items()},
                     "bound": 1e-5,
                     "passed": bool(all(rel[k] <= 1e-5 for k, (_, _, g) in fd.items() if g))},
        "calibration": {"start": {"mu_s": base["mu_s"], "E": base["E"]},
                        "target": {"mu_s": workloads.ADJOINT_TARGET["mu_s"],
                                   "E": workloads.ADJOINT_TARGET["E"]},
                        "found": {"mu_s": cal["mu_s"], "E": cal["E"]},
                        "iterations": cal["iterations"], "residual": cal["residual"],
                        "seconds": cal_s,
                        "how": "Gauss-Newton on the three residuals, Jacobian from three "
                               "reverse sweeps per iteration (calibrate.py)"},
        "roofline": None,
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        log("bench: CPU baseline (oracle forward + numpy reverse sweep, 20 steps)")
        import adjoint_oracle
        nsteps = 20
        t0 = time.perf_counter()
        recs, fin = adjoint_oracle.record(w, nsteps, num_threads=1)
        _, seed20 = calibrate.loss_and_seed(fin["x"], fin["y"], fin["active"], target)
        lam0, gcpu = adjoint_oracle.adjoint(recs, seed20, w["material"], w["N"], w["h"],
                                            w["dt"])
        el = time.perf_counter() - t0
        # the GPU reverse sweep of the same 20 steps, against it
        a20 = pympm.Adjoint(cfg_for(w["material"]), w["particles"], checkpoint_every=20)
        a20.forward(nsteps)
        gg, ggi = a20.backward(seed20)
        a20.close()
        cmp = {k: abs(gg[k] - gcpu[adjoint_oracle.PARAMS.index(k)])
               / abs(gcpu[adjoint_oracle.PARAMS.index(k)]) for k in ("mu_s", "E")}
        for k in ("vx", "vy"):
            cmp[k] = float(np.abs(ggi[k] - lam0[k]).max() / np.abs(lam0[k]).max())
        cmp = {k: float(v) for k, v in cmp.items()}
        result["accuracy"]["cpu_adjoint_20_steps"] = {"rel_err": cmp, "bound": 1e-9,
                                                      "passed": bool(max(cmp.values()) <= 1e-9)}
        result["accuracy"]["passed"] = bool(result["accuracy"]["passed"]
                                            and max(cmp.values()) <= 1e-9)
        result["cpu_baseline"] = {
            "value": npart * nsteps / el, "unit": "particle-steps/s (forward + reverse)",
            "cores": 1, "kind": "port",
            "sample": (f"the first {nsteps} of the 200 steps ({el:.1f} s): oracle/usl_oracle.c "
                       "forward on 1 thread, recorded step by step, then the numpy reverse "
                       "sweep of oracle/adjoint_oracle.py")}
        result["speedup_vs_cpu"] = result["value"] / result["cpu_baseline"]["value"]
    if emit:
        print(json.dumps(result), flush=True)
    return result


def cpu_baseline_members(sample, budget_s):
    """The CPU restatement over a sample of ensemble members, each stepped
    for an equal share of `budget_s` on all host cores."""
    import pyoracle
    flags = open("/proc/cpuinfo").read() if os.path.exists("/proc/cpuinfo") else ""
    variant = "v4" if " avx512f" in flags else "v3"
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    psteps, el_sum, steps_sum = 0, 0.0, 0
    for w in sample:
        o = pyoracle.from_workload(w, variant=variant, num_threads=threads)
        o.run(1)
        steps, t0 = 0, time.perf_counter()
        while True:
            o.run(1)
            steps += 1
            el = time.perf_counter() - t0
            if el >= budget_s / len(sample) or steps >= 2000:
                break
        psteps += int(o.particles()["active"].sum()) * steps
        el_sum += el
        steps_sum += steps
        o.close()
    return {"value": psteps / el_sum, "unit": "particle-steps/s", "cores": threads,
            "kind": "port",
            "sample": (f"{len(sample)} members (every 32nd: "
                       f"{[len(w['particles']['x']) for w in sample]} particles), "
                       f"{steps_sum} steps in all ({el_sum:.1f} s), oracle/usl_oracle.c "
                       f"x86-64-{variant} build, {threads} pthreads")}


def copy_probe():
    """A 256 MiB device-to-device copy through the HIP runtime, five times:
    the known-size transfer that calibrates rocprofv3's FETCH_SIZE and
    WRITE_SIZE counters (bytes per KiB reported) for `roofline.traffic`."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = 256 * 1024 * 1024
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) or \
            hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)):
        sys.exit("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(n))
    for _ in range(5):
        if hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3):   # hipMemcpyDeviceToDevice
            sys.exit("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    print(json.dumps({"copy_probe_bytes": n, "copies": 5}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the 256 MiB copy that calibrates the HBM traffic counters")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true",
                    help="skip the CPU-baseline leg (and with it the accuracy gate)")
    ap.add_argument("--no-subrecords", action="store_true",
                    help="silo only: no batched-ensemble, adjoint and nonlocal-heap sub-records")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--repeats", type=int, default=3, help="adjoint workload: timed runs")
    ap.add_argument("--workload", choices=["silo", "heap", "ensemble", "adjoint"],
                    default="silo",
                    help="silo: the plain USL step (default, the headline number); "
                         "heap: the g_nonlocal fluidity heap; "
                         "ensemble: 256 uneven members stepped as one batch; "
                         "adjoint: forward + reverse sweep of a 200-step run")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the USL step is built and timed on a single "
              "MI355X; multi-GPU sharding is a design note, not a measured path.",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "heap":
        bench_heap(args)
        return
    if args.workload == "ensemble":
        bench_ensemble(args)
        return
    if args.workload == "adjoint":
        bench_adjoint(args)
        return
    result = bench_silo(args)
    if not args.no_subrecords:
        sub = argparse.Namespace(**vars(args))
        sub.steps, sub.warmup, sub.no_cpu_baseline, sub.repeats = 20, 5, True, 1
        log("bench: sub-record: batched ensemble against the same members as plain steps")
        ens = bench_ensemble(sub, emit=False)
        result["batched_vs_plain"] = ens["batched_vs_plain"]
        result["batched"] = {k: ens[k] for k in ("value", "unit", "ms_per_step", "config",
                                                 "roofline", "plain", "accuracy")}
        log("bench: sub-record: adjoint (forward + reverse) against the plain forward run")
        sub.no_cpu_baseline = args.no_cpu_baseline   # its CPU leg also checks the reverse sweep
        adj = bench_adjoint(sub, emit=False)
        result["adjoint_vs_plain"] = adj["adjoint_vs_plain"]
        result["adjoint"] = {k: adj[k] for k in ("value", "unit", "forward_ms", "reverse_ms",
                                                 "plain_forward_ms", "config", "accuracy",
                                                 "calibration", "cpu_baseline")}
        log("bench: sub-record: the g_nonlocal fluidity heap")
        sub.steps, sub.warmup, sub.no_cpu_baseline = 100, 20, True
        heap = bench_heap(sub, emit=False)
        result["nonlocal"] = {k: heap[k] for k in ("value", "unit", "ms_per_step", "config",
                                                   "cg_iterations_mean", "cg_iterations_max",
                                                   "cg_final_rel_residual")}
    print(json.dumps(result), flush=True)


def bench_silo(args):
    """The headline: the square silo, state resident in HBM.  Pass A times
    `steps` steps between two HIP events and nothing else (value,
    ms_per_step); pass B runs the same number of steps with each kernel's
    dispatch recording its own start and end events, for the per-kernel
    durations the roofline uses.  Wall time
    around pass A is kept as a cross-check.  In the CPU-baseline leg the
    accuracy gate compares a separate GPU run of the same workload with the
    oracle's parity build after 1 and after 100 steps (bounds 1e-12 and 1e-8)
    before the CPU restatement is timed."""
    w = workloads.silo()
    npart = len(w["particles"]["x"])
    nnodes = w["N"] * w["N"]
    log(f"bench: {w['name']} particles={npart} nodes={nnodes} warmup={args.warmup} "
        f"steps={args.steps}")
    cfg = pympm.make_config(w["N"], w["h"], w["dt"], material=w["material"],
                            bc_kind=w["bc_kind"], bc_fp64=w["bc_fp64"], bc_int=w["bc_int"],
                            g_mag=w["g_mag"], ramp_time=w["ramp_time"])
    sim = pympm.Sim(cfg, w["particles"])
    sim.step(args.warmup)
    sim.sync()
    nact0 = int(sim.active().sum())

    t0 = time.perf_counter()
    span_ms = sim.time_span(args.steps)       # pass A: two events, n steps
    wall = time.perf_counter() - t0
    nact1 = int(sim.active().sum())
    ms = sim.time_steps(args.steps)           # pass B: per-kernel dispatch events
    sim.close()
    pstep = (nact0 + nact1) / 2.0 * args.steps
    step_ms = span_ms / args.steps

    p2g_ms, g2p_ms = ms[1] / args.steps, ms[2] / args.steps
    p2g_bytes = P2G_BYTES_PER_PARTICLE * npart + P2G_BYTES_PER_NODE * nnodes
    g2p_bytes = G2P_BYTES_PER_PARTICLE * npart + G2P_BYTES_PER_NODE * nnodes
    kernels = {
        "k_p2g_grid": (p2g_ms, p2g_bytes),
        "k_g2p_stress": (g2p_ms, g2p_bytes),
    }
    dom = max(kernels, key=lambda k: kernels[k][0])
    dom_ms, dom_bytes = kernels[dom]
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    step_bytes = p2g_bytes + g2p_bytes
    result = {
        "metric": "particle-steps/sec FP64 USL step, % HBM roofline; batched+adjoint vs plain",
        "value": pstep / (span_ms * 1e-3),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_ms,
        "ms_per_step_wall": wall * 1e3 / args.steps,
        "ms_per_step_kernel_sum": (ms[1] + ms[2]) / args.steps,
        "ms_per_step_kernel_pass": ms[0] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded lattice + jitter, generated in bench.py)",
        "config": {"workload": "square-silo discharge, 512x512 cells (513^2 nodes), "
                               "1,048,576 particles, g_local_mu2, dt=2e-6",
                   "particles": npart, "nodes": nnodes, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": read_traffic(dom),
                     "traffic_GBs": _traffic_rate(read_traffic(dom), dom_ms), "kernel": dom,
                     "kernel_ms": dom_ms, "algorithmic_bytes_per_launch": dom_bytes},
        "kernels_ms": {k: v[0] for k, v in kernels.items()},
        "step_roofline": {"algorithmic_bytes": step_bytes,
                          "achieved_GBs": step_bytes / (step_ms * 1e-3) / 1e9,
                          "frac": step_bytes / (step_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "accuracy": None,
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        log("bench: CPU-baseline leg: accuracy gate against the oracle's parity build")
        result["accuracy"] = accuracy_gate(w, cfg)
        if not result["accuracy"]["passed"]:
            print(json.dumps(result), flush=True)
            sys.exit("bench: the accuracy gate failed; the throughput does not count")
        log("bench: CPU-baseline leg: oracle restatement on all host cores")
        result["cpu_baseline"] = cpu_baseline(w, args.cpu_budget)
        result["speedup_vs_cpu"] = result["value"] / result["cpu_baseline"]["value"]
    return result


GATE_FIELDS = ("x", "y", "vx", "vy", "txx", "txy", "tyy")


def accuracy_gate(w, cfg, steps=(1, 100)):
    """Per-field relative error max|gpu - cpu| / max|cpu| of position,
    velocity and stress after steps[0] (bound 1e-12) and steps[1] (bound
    1e-8), the GPU run against the oracle's parity build (-ffp-contract=off,
    portable exp) on identical inputs.  Part of the CPU-baseline leg."""
    import pyoracle
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    sim = pympm.Sim(cfg, w["particles"])
    o = pyoracle.from_workload(w, num_threads=threads)
    done, out = 0, {"steps": list(steps), "bounds": [1e-12, 1e-8], "max_rel_err": [],
                    "bitwise": [], "fields": list(GATE_FIELDS), "checker": "oracle parity build",
                    "threads": threads}
    for n in steps:
        sim.step(n - done)
        o.run(n - done)
        done = n
        ref = o.particles()
        err, same = 0.0, True
        for k in GATE_FIELDS:
            got = sim.get(k)
            scale = float(np.abs(ref[k]).max()) or 1.0
            err = max(err, float(np.abs(got - ref[k]).max()) / scale)
            same = same and bool(np.array_equal(got, ref[k]))
        out["max_rel_err"].append(err)
        out["bitwise"].append(same)
    sim.close()
    o.close()
    out["passed"] = all(e <= b for e, b in zip(out["max_rel_err"], out["bounds"]))
    return out


if __name__ == "__main__":
    main()
