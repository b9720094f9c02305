#!/usr/bin/env python3
"""Benchmark: leaf-cell updates/s of the FP64 octree leaf update on one MI355X.

Workload: 3D single-phase Euler in a periodic unit box on a uniform octree at
level 8 (256^3 = 16,777,216 leaves), smoothed random density and a
divergence-free velocity field of RMS Mach 0.5, p = 1, gamma = 1.4, LTS on,
no refinement (HLLC-type flux + MUSCL-Hancock, minmod).  A step is one
global step of integrate(); a leaf-cell update is one leaf advanced by one
LTS substep.  Inputs are resident in HBM before the timed region.

Prints one JSON line (see the keys below).  `--gpus N` with N > 1 exits: this
engine runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

B_ALG = 8 * (2 * 5 + 2)  # bytes per leaf-cell update: state in + out, cell time in + out (euler<3>: NE = 5)
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# 1024 SIMDs x 2.4 GHz, one wave64 VALU instruction per 4 cycles, 64 lanes
VALU_PEAK_TLANE = 1024 * 2.4e9 / 4 * 64 / 1e12


def profile_figures():
    """Per-update figures of the box workload from the committed rocprofv3
    summaries: measured HBM bytes (FETCH_SIZE / WRITE_SIZE, calibrated on a
    1 GiB device copy), FP64 FLOPs and VALU instructions per leaf-cell update.
    None where a summary is missing."""
    out = {"bytes": None, "flops": None, "valu": None}
    prof = os.path.join(ROOT, "profiles")
    try:
        with open(os.path.join(prof, "leaf_update_euler3d_256_hbm_bytes.txt")) as f:
            for line in f:
                if line.startswith("Measured HBM bytes per leaf-cell update:"):
                    out["bytes"] = float(line.split(":")[1].split()[0])
        with open(os.path.join(prof, "leaf_update_euler3d_256_valu.txt")) as f:
            for line in f:
                if "FP64 FLOPs per leaf-cell update" in line:
                    out["flops"] = float(line.rsplit(":", 1)[1])
                if "VALU instructions per leaf-cell update" in line:
                    out["valu"] = float(line.rsplit(":", 1)[1])
    except OSError:
        pass
    return out


B_ALG_MULTI2 = 8 * (2 * 6 + 2) + 16  # multi<2>: state in + out, time in + out, key/level/flags
B_ALG_MULTI3 = 8 * (2 * 7 + 2) + 16  # multi<3>


def substep_roofline(b_alg, updates, launches, kern_ms, kernel):
    """HBM roofline of one level substep (all its kernels) from the HIP-event
    kernel time: algorithmic bytes per launch / average launch duration."""
    avg_ms = kern_ms / max(launches, 1)
    achieved = b_alg * updates / max(launches, 1) / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None, "bytes_per_update": b_alg, "kernel": kernel,
            "avg_launch_ms": avg_ms, "launches": launches}


def oracle_instances_baseline(fam, trees, seconds=20.0):
    """The oracle (OpenMP) stepping the batch's instances one after another
    from the same initial trees, one global step each, until about `seconds`
    of CPU work: a bounded sample of the batched workload."""
    from oracle.binding import Oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    n = k = 0
    el = 0.0
    for (case, _), tree in zip(fam, trees):
        o = Oracle(case, tree)
        t0 = time.perf_counter()
        o.integrate()
        el += time.perf_counter() - t0
        n += o.cell_updates
        k += 1
        o.close()
        if el >= seconds:
            break
    return {"value": n / el, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"the first global step of the first {k} instances from the same initial states "
                      f"({n} leaf-cell updates in {el:.1f} s, OpenMP {threads} threads)"}


def cpu_baseline(case, tree, seconds_target=20.0):
    """The oracle (C restatement, OpenMP) on the same initial tree: one global step."""
    from oracle.binding import Oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = Oracle(case, tree)
    u0 = o.cell_updates
    t0 = time.perf_counter()
    o.integrate()
    el = time.perf_counter() - t0
    n = o.cell_updates - u0
    o.close()
    return {"value": n / el, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"one global step of the same 256^3 workload from the same initial state "
                      f"({n} leaf-cell updates in {el:.1f} s, OpenMP {threads} threads)"}


def oracle_steps_baseline(case, tree, max_steps, seconds=10.0):
    """The oracle (OpenMP) stepping one case from its initial tree, global
    step after global step (remeshing in every substep), until about
    `seconds` of CPU work or `max_steps` steps."""
    from oracle.binding import Oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = Oracle(case, tree)
    t1 = time.perf_counter()
    n = 0
    while time.perf_counter() - t1 < seconds and n < max_steps:
        o.integrate()
        n += 1
    el = time.perf_counter() - t1
    out = {"value": o.cell_updates / el, "unit": "cell-updates/s", "cores": threads, "kind": "port",
           "sample": f"the first {n} global steps of the held-out case from the same initial state, remeshing in "
                     f"every substep ({o.cell_updates} updates in {el:.1f} s, OpenMP {threads} threads)"}
    o.close()
    return out


def mlp_roofline(gemm_tflops):
    """MFMA roofline of the predictor's training: FLOPs of its GEMMs over the
    training time, against the dense BF16 peak."""
    return {"bound": "mfma", "achieved": gemm_tflops, "peak": 2500.0, "unit": "TFLOP/s",
            "frac": gemm_tflops / 2500.0, "traffic": None,
            "kernel": "MLP training (BF16 MFMA products, FP32 accumulation, whole training time)"}


def copy_probe(gib=1.0):
    """A plain device-to-device copy of known size (HIP runtime), run in the
    same process so hardware byte counters can be calibrated against it."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = int(gib * (1 << 30))
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    hip.hipMemset(a, 1, C.c_size_t(n))
    hip.hipDeviceSynchronize()
    hip.hipMemcpy(b, a, C.c_size_t(n), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(f"[bench] copy probe: {n} bytes device to device", file=sys.stderr, flush=True)


def batched(args):
    """64 independent 2D shock-bubble instances (five-equation model, seeded
    Mach, radius, density ratio and refinement depth), stepped one integrator
    at a time ("plain") and as one batch (one launch per op and substep for
    all instances, autotuned launch, graph replay).  Both start from the same
    states and time the same global steps, so they do the same cell updates."""
    from dacti_amd import scene, workloads
    from dacti_amd.integrator import Batch, Integrator
    fam = workloads.shock_bubble_family(args.instances, seed=args.seed)
    trees = [scene.build_tree(c) for c, _ in fam]
    print(f"[bench] {len(trees)} instance trees, {sum(len(t.leaves()) for t in trees)} leaves", file=sys.stderr,
          flush=True)
    warm = max(args.warmup, 6)

    # an instance whose state turns NaN stops (force_break), in both forms
    plain = [Integrator(c, t) for (c, _), t in zip(fam, trees)]
    alive = [True] * len(plain)

    def plain_step():
        for k, g in enumerate(plain):
            if alive[k] and g.integrate()[1]:
                alive[k] = False

    tw = time.perf_counter()
    for w in range(warm):
        plain_step()
        print(f"[bench] plain warmup step {w} done at {time.perf_counter() - tw:.1f} s", file=sys.stderr, flush=True)
    u0 = sum(g.stats().cell_updates for g in plain)
    t0 = time.perf_counter()
    for k in range(args.steps):
        plain_step()
        print(f"[bench] plain step {k} done", file=sys.stderr, flush=True)
    el_plain = time.perf_counter() - t0
    up_plain = sum(g.stats().cell_updates for g in plain) - u0
    print(f"[bench] plain: {args.steps} steps in {el_plain:.3f} s", file=sys.stderr, flush=True)

    members = [Integrator(c, t) for (c, _), t in zip(fam, trees)]
    b = Batch(members)
    if not args.no_tune:
        b.autotune()
    for w in range(warm):
        b.integrate()
        print(f"[bench] batched warmup/tuning step {w} done", file=sys.stderr, flush=True)
    tuned = b.tune_results()
    b.set_timing(True)
    s0 = b.stats()
    t0 = time.perf_counter()
    for k in range(args.steps):
        b.integrate()
        print(f"[bench] batched step {k} done", file=sys.stderr, flush=True)
    el = time.perf_counter() - t0
    s1 = b.stats()
    up = s1.cell_updates - s0.cell_updates
    print(f"[bench] batched: {args.steps} steps in {el:.3f} s", file=sys.stderr, flush=True)
    same = all(np.array_equal(a.get_nodes()["u"], m.get_nodes()["u"], equal_nan=True)
               for a, m in zip(plain, members))
    blk, grd, gr = b.launch()
    roof = substep_roofline(B_ALG_MULTI2, up, s1.flux_launches - s0.flux_launches,
                            s1.substep_kernel_ms - s0.substep_kernel_ms,
                            "one level substep of all instances (head, remesh and tail sequences)")
    base = None
    if not args.no_cpu_baseline:
        base = oracle_instances_baseline(fam, trees)
    out = {
        "metric": "cell-updates/s (HLLC+MUSCL FP64) & %HBM roofline; learned/batched vs plain",
        "value": up / el,
        "unit": "cell-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": warm,
        "ms_per_step": el * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{len(fam)} independent 2D shock-bubble instances (five-equation model), "
                               "seeded Mach U[1.1,3], radius U[0.1,0.3], density ratio log-U[0.1,10], "
                               "4-7 levels above a 64x32 base, AMR and LTS on",
                   "instances": len(fam), "leaves_initial": sum(len(t.leaves()) for t in trees),
                   "cell_updates_timed": up, "launch": {"block": blk, "grid": grd, "graphs": gr},
                   "autotune": [{"block": r[0], "grid": r[1], "graphs": r[2], "ms": r[3]} for r in tuned]},
        "plain": {"value": up_plain / el_plain, "cell_updates_timed": up_plain,
                  "ms_per_step": el_plain * 1e3 / args.steps},
        "batched_over_plain": (up / el) / (up_plain / el_plain),
        "same_state_as_plain": bool(same),
        "instances_stopped_nan": int(len(plain) - sum(alive)),
        "roofline": roof,
        "cpu_baseline": base,
    }
    b.close()
    for g in plain + members:
        g.close()
    print(json.dumps(out), flush=True)


def learned(args):
    """Learned refinement predictor on 2D shock-bubble runs (five-equation
    model): solver runs generate features and labels (the mesh `horizon`
    global steps later), the BF16-MFMA MLP is trained on the GPU, and a
    held-out case is run three ways from the same initial state: remeshing
    in every substep (the reference's behaviour), remeshing every 4 global
    steps with the estimator alone, and every `horizon` steps with the
    predictor refining ahead (estimator as fallback)."""
    from dacti_amd import predictor, scene, workloads
    from dacti_amd.mlp import MLP

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    h = args.horizon
    fam = workloads.shock_bubble_family(args.train_runs + 1, seed=args.seed + 100, base=5, extra_range=(3, 3),
                                        mach_range=(1.2, 2.2), ratio_range=(0.2, 5.0))
    Xs, ys, gs = [], [], []
    t0 = time.perf_counter()
    for k, (case, prm) in enumerate(fam[:-1]):
        for X, y, gi in predictor.run_samples(case, args.train_steps, h, window=True):
            Xs.append(X), ys.append(y), gs.append(gi)
        log(f"training run {k} done ({sum(len(y) for y in ys)} samples, {time.perf_counter() - t0:.0f} s)")
    X, y = np.concatenate(Xs), np.concatenate(ys)
    std = predictor.Standardizer(X)
    m = MLP(X.shape[1], 128, 3, 3, seed=args.seed)
    loss = m.train(std(X), y, epochs=args.epochs, batch=1024, lr=2e-3, seed=1)
    ms, fl = m.train_stats()
    log(f"trained: loss {loss:.3f}, {fl / ms / 1e9:.2f} TFLOP/s")
    case, prm = fam[-1]
    tree = scene.build_tree(case)
    held = predictor.run_samples(case, args.train_steps, h, tree=tree, window=True)
    Xt = np.concatenate([a for a, _, _ in held])
    yt = np.concatenate([b for _, b, _ in held])
    gt = np.concatenate([c for _, _, c in held])
    pred = m.predict(std(Xt))
    miss_net, miss_ind = predictor.missed_refinement(pred, yt), predictor.missed_refinement(gt, yt)
    log(f"held-out: missed refinement net {miss_net:.3f}, indicator {miss_ind:.3f}")
    ref, t_ref, u_ref, _ = predictor.run_cadence(case, tree, args.steps, 1)
    log(f"reference cadence run: {t_ref:.1f} s")
    p4, t_p4, u_p4, _ = predictor.run_cadence(case, tree, args.steps, 4)
    log(f"estimator every 4 steps: {t_p4:.1f} s")
    ln, t_ln, u_ln, flagged = predictor.run_cadence(case, tree, args.steps, h, model=m, std=std, rings=args.rings)
    log(f"learned every {h} steps: {t_ln:.1f} s")
    out = {
        "metric": "cell-updates/s (HLLC+MUSCL FP64) & %HBM roofline; learned/batched vs plain",
        "value": u_ln / t_ln,
        "unit": "cell-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": 0,
        "ms_per_step": t_ln * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"2D shock-bubble (five-equation model), 32x16 base + 3 levels, learned refinement "
                               f"every {h} global steps vs the estimator every 4 and every substep",
                   "held_out_case": prm, "train_runs": args.train_runs, "train_samples": int(len(y)),
                   "features": int(X.shape[1]), "mlp": "3x128 ReLU, BF16 MFMA, FP32 accumulation",
                   "horizon_steps": h, "labels": "finer at any of the next horizon steps",
                   "dilation_rings": args.rings},
        "learned": {"cell_updates_per_s": u_ln / t_ln, "seconds": t_ln, "cell_updates": u_ln,
                    "flagged_leaves": flagged, "rel_l1_vs_every_substep": predictor.relative_l1(ref, ln),
                    "missed_refinement": miss_net, "accuracy": float((pred == yt).mean())},
        "plain_every_4": {"cell_updates_per_s": u_p4 / t_p4, "seconds": t_p4, "cell_updates": u_p4,
                          "rel_l1_vs_every_substep": predictor.relative_l1(ref, p4),
                          "missed_refinement_of_present_flags": miss_ind},
        "every_substep": {"cell_updates_per_s": u_ref / t_ref, "seconds": t_ref, "cell_updates": u_ref},
        "training": {"loss": loss, "ms": ms, "gemm_tflops": fl / ms / 1e9},
        # the same simulated time in every run: wall-time ratios are time-to-solution speedups
        "learned_speedup_vs_every_4": t_p4 / t_ln,
        "learned_speedup_vs_every_substep": t_ref / t_ln,
        "roofline": mlp_roofline(fl / ms / 1e9),
        "cpu_baseline": None if args.no_cpu_baseline else oracle_steps_baseline(case, tree, args.steps),
    }
    m.close()
    print(json.dumps(out), flush=True)


def learned3d(args):
    """Learned refinement predictor on 3D shock-drop runs (five-equation
    stiffened gas, config 5): seeded Mach numbers uniform in [1.5, 3.5],
    features (3x3x3 neighbourhood primitives + gradients), labels (the mesh
    is finer at any of the next `horizon` global steps), inference and split
    flags all computed in device memory; the BF16-MFMA MLP trains on the GPU;
    a held-out Mach number is run remeshing in every substep (the reference),
    with the estimator every 4 global steps, and every `every` steps with the
    predictor refining ahead (estimator as fallback)."""
    from dacti_amd import predictor, scene, workloads
    from dacti_amd.devmem import DeviceArray
    from dacti_amd.mlp import MLP

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    h, every = args.horizon, args.every
    rng = np.random.default_rng(args.seed + 500)
    machs = [float(x) for x in rng.uniform(1.5, 3.5, args.train_runs + 1)]
    Xs, ys, gs = [], [], []
    t0 = time.perf_counter()
    for k, mach in enumerate(machs[:-1]):
        case = workloads.shock_drop(base=args.base, extra=args.extra, mach=mach, max_dt=args.drop_max_dt)
        for X, y, gi in predictor.gpu_samples(case, args.train_steps, h):
            Xs.append(X), ys.append(y), gs.append(gi)
        log(f"training run {k} (Mach {mach:.2f}) done: {sum(len(y) for y in ys)} samples, "
            f"{time.perf_counter() - t0:.0f} s")
    X, y = np.concatenate(Xs), np.concatenate(ys)
    del Xs, ys
    std = predictor.Standardizer(X)
    m = MLP(X.shape[1], 128, 3, 3, seed=args.seed)
    loss = m.train(std(X), y, epochs=args.epochs, batch=1024, lr=2e-3, seed=1)
    ms, fl = m.train_stats()
    log(f"trained on {len(y)} samples x {X.shape[1]} features: loss {loss:.3f}, {fl / ms / 1e9:.2f} TFLOP/s")
    case = workloads.shock_drop(base=args.base, extra=args.extra, mach=machs[-1], max_dt=args.drop_max_dt)
    tree = scene.build_tree(case)
    held = predictor.gpu_samples(case, args.train_steps, h, tree=tree)
    Xt = np.concatenate([a for a, _, _ in held])
    yt = np.concatenate([b for _, b, _ in held])
    gt = np.concatenate([c for _, _, c in held])
    pred = m.predict(std(Xt))
    miss_net, miss_ind = predictor.missed_refinement(pred, yt), predictor.missed_refinement(gt, yt)
    log(f"held-out Mach {machs[-1]:.2f}: missed refinement net {miss_net:.3f}, indicator {miss_ind:.3f}")
    mu = DeviceArray.from_numpy(std.mu.astype(np.float32))
    sd = DeviceArray.from_numpy(std.sd.astype(np.float32))
    ref, t_ref, u_ref, _ = predictor.run_cadence_gpu(case, tree, args.steps, 1)
    log(f"reference cadence: {t_ref:.2f} s")
    p4, t_p4, u_p4, _ = predictor.run_cadence_gpu(case, tree, args.steps, 4)
    log(f"estimator every 4 steps: {t_p4:.2f} s")
    ln, t_ln, u_ln, flagged = predictor.run_cadence_gpu(case, tree, args.steps, every, model=m, mu=mu, sd=sd,
                                                        rings=args.rings)
    log(f"learned every {every} steps: {t_ln:.2f} s")
    sweep = []
    for spec in (args.cadences.split(",") if args.cadences else []):
        ev, rg = (int(v) for v in spec.split(":"))
        sn, st, su, sf = predictor.run_cadence_gpu(case, tree, args.steps, ev, model=m, mu=mu, sd=sd, rings=rg)
        p_ev, t_ev, u_ev, _ = predictor.run_cadence_gpu(case, tree, args.steps, ev)
        sweep.append({"every": ev, "rings": rg, "learned_seconds": st, "learned_cell_updates": su,
                      "learned_rel_l1": predictor.relative_l1(ref, sn), "estimator_seconds": t_ev,
                      "estimator_rel_l1": predictor.relative_l1(ref, p_ev), "flagged": sf})
        log(f"sweep every {ev} rings {rg}: learned {st:.2f} s L1 {sweep[-1]['learned_rel_l1']:.4f}, "
            f"estimator {t_ev:.2f} s L1 {sweep[-1]['estimator_rel_l1']:.4f}")
    out = {
        "metric": "cell-updates/s (HLLC+MUSCL FP64) & %HBM roofline; learned/batched vs plain",
        "value": u_ln / t_ln,
        "unit": "cell-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": 0,
        "ms_per_step": t_ln * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"3D shock-drop (five-equation stiffened gas), octree levels {args.base}-"
                               f"{args.base + args.extra}, learned refinement every {every} global steps vs the "
                               f"estimator every 4 and every substep",
                   "train_machs": machs[:-1], "held_out_mach": machs[-1], "train_samples": int(len(y)),
                   "features": int(X.shape[1]), "mlp": "3x128 ReLU, BF16 MFMA, FP32 accumulation",
                   "horizon_steps": h, "labels": "finer at any of the next horizon global steps",
                   "dilation_rings": args.rings, "features_and_inference": "device memory",
                   "max_dt": args.drop_max_dt},
        "learned": {"cell_updates_per_s": u_ln / t_ln, "seconds": t_ln, "cell_updates": u_ln,
                    "flagged_leaves": flagged, "rel_l1_vs_every_substep": predictor.relative_l1(ref, ln),
                    "missed_refinement": miss_net, "accuracy": float((pred == yt).mean())},
        "plain_every_4": {"cell_updates_per_s": u_p4 / t_p4, "seconds": t_p4, "cell_updates": u_p4,
                          "rel_l1_vs_every_substep": predictor.relative_l1(ref, p4),
                          "missed_refinement_of_present_flags": miss_ind},
        "every_substep": {"cell_updates_per_s": u_ref / t_ref, "seconds": t_ref, "cell_updates": u_ref},
        "training": {"loss": loss, "ms": ms, "gemm_tflops": fl / ms / 1e9},
        "learned_speedup_vs_every_4": t_p4 / t_ln,
        "learned_speedup_vs_every_substep": t_ref / t_ln,
        "cadence_sweep": sweep,
        "roofline": mlp_roofline(fl / ms / 1e9),
        "cpu_baseline": None if args.no_cpu_baseline else oracle_steps_baseline(case, tree, args.steps),
    }
    m.close()
    print(json.dumps(out), flush=True)


def bubble(args):
    """2D shock-bubble (five-equation model): 64x32 base quadtree + `--extra`
    levels, Mach 1.22 shock in air hitting a helium-like bubble (gamma 1.67,
    density 0.138, radius 0.25), HLLC + MUSCL minmod, CFL 0.5, LTS on,
    remeshing in every substep: the plain method, GPU vs the CPU port."""
    from dacti_amd import scene, workloads
    from dacti_amd.integrator import Integrator
    case = workloads.shock_bubble(base=6, extra=args.extra)
    tree = scene.build_tree(case)
    print(f"[bench] shock-bubble tree: {tree.n} nodes, {len(tree.leaves())} leaves", file=sys.stderr, flush=True)
    g = Integrator(case, tree)
    if args.graphs or args.grid:
        # a fixed grid lets each substep run planned on the device (one wait)
        g.set_launch(block=args.block, grid=args.grid or 2048, graphs=bool(args.graphs))
    for w in range(args.warmup):
        g.integrate()
    g.set_timing(True)
    s0 = g.stats()
    t0 = time.perf_counter()
    for k in range(args.steps):
        if g.integrate()[1]:
            raise RuntimeError("NaN state (force_break)")
        if k % 10 == 9:
            print(f"[bench] step {k} done", file=sys.stderr, flush=True)
    el = time.perf_counter() - t0
    s1 = g.stats()
    updates = s1.cell_updates - s0.cell_updates
    base = None
    if not args.no_cpu_baseline:
        from oracle.binding import Oracle
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        o = Oracle(case, tree)
        t1 = time.perf_counter()
        n_cpu = 0
        while time.perf_counter() - t1 < 10.0 and n_cpu < args.warmup + args.steps:
            o.integrate()
            n_cpu += 1
        el_cpu = time.perf_counter() - t1
        base = {"value": o.cell_updates / el_cpu, "unit": "cell-updates/s", "cores": threads, "kind": "port",
                "sample": f"the first {n_cpu} global steps from the same initial state ({o.cell_updates} updates in "
                          f"{el_cpu:.1f} s, OpenMP {threads} threads)"}
        o.close()
    out = {
        "metric": "cell-updates/s (HLLC+MUSCL FP64) & %HBM roofline; learned/batched vs plain",
        "value": updates / el, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": el * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"2D shock-bubble (five-equation model), 64x32 base + {args.extra} levels, "
                               "remesh every substep, LTS on",
                   "nodes": s1.n_nodes, "active_leaves": s1.n_active, "max_acti_level": s1.max_acti_level,
                   "cell_updates_timed": updates, "graphs": bool(args.graphs),
                   "launch": {"block": args.block if (args.graphs or args.grid) else 256,
                              "grid": (args.grid or 2048) if (args.graphs or args.grid) else 0}},
        "roofline": substep_roofline(B_ALG_MULTI2, updates, s1.flux_launches - s0.flux_launches,
                                     s1.substep_kernel_ms - s0.substep_kernel_ms,
                                     "one level substep (head, remesh and tail sequences)"),
        "cpu_baseline": base,
    }
    g.close()
    print(json.dumps(out), flush=True)


def drop(args):
    """3D shock-drop (five-equation stiffened gas, SI units): Mach 2.4 shock
    in air hitting a 1 mm water drop in a 16 x 8 x 8 mm box, octree levels
    `--base` to `--base + --extra`, remeshing in every substep, LTS on."""
    from dacti_amd import scene, workloads
    from dacti_amd.integrator import Integrator
    case = workloads.shock_drop(base=args.base, extra=args.extra)
    tree = scene.build_tree(case)
    print(f"[bench] shock-drop tree: {tree.n} nodes, {len(tree.leaves())} leaves", file=sys.stderr, flush=True)
    g = Integrator(case, tree)
    for w in range(args.warmup):
        g.integrate()
        print(f"[bench] warmup step {w} done", file=sys.stderr, flush=True)
    g.set_timing(True)
    s0 = g.stats()
    t0 = time.perf_counter()
    for k in range(args.steps):
        if g.integrate()[1]:
            raise RuntimeError("NaN state (force_break)")
        print(f"[bench] step {k} done", file=sys.stderr, flush=True)
    el = time.perf_counter() - t0
    s1 = g.stats()
    updates = s1.cell_updates - s0.cell_updates
    launches = s1.flux_launches - s0.flux_launches
    roof = substep_roofline(B_ALG_MULTI3, updates, launches, s1.substep_kernel_ms - s0.substep_kernel_ms,
                            "one level substep (head, remesh and tail sequences)")
    base = None
    if not args.no_cpu_baseline:
        from oracle.binding import Oracle
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        o = Oracle(case, tree)
        t1 = time.perf_counter()
        o.integrate()
        el_cpu = time.perf_counter() - t1
        base = {"value": o.cell_updates / el_cpu, "unit": "cell-updates/s", "cores": threads, "kind": "port",
                "sample": f"the first global step from the same initial state ({o.cell_updates} updates in "
                          f"{el_cpu:.1f} s, OpenMP {threads} threads)"}
        o.close()
    out = {
        "metric": "cell-updates/s (HLLC+MUSCL FP64) & %HBM roofline; learned/batched vs plain",
        "value": updates / el, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": el * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"3D shock-drop (five-equation stiffened gas), octree levels {args.base}-"
                               f"{args.base + args.extra}, remesh every substep, LTS on",
                   "nodes": s1.n_nodes, "active_leaves": s1.n_active, "max_acti_level": s1.max_acti_level,
                   "cell_updates_timed": updates, "substeps_timed": launches},
        "roofline": roof,
        "cpu_baseline": base,
    }
    g.close()
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed global steps (default 20; 48 for learned, 4 for batched; 3 for the shock-drop, whose state "
                         "turns NaN in step 6)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", action="store_true", help="run a 1 GiB device copy first (counter calibration)")
    ap.add_argument("--workload", choices=("box", "batched", "learned", "learned3d", "drop", "bubble"), default="box",
                    help="box: 3D periodic Euler box (default); batched: shock-bubble instances, batch vs plain; "
                         "learned: refinement predictor vs the estimator; drop: 3D shock-drop with AMR; "
                         "bubble: 2D shock-bubble with AMR")
    ap.add_argument("--graphs", action="store_true", help="box, bubble: persistent grid and graph replay")
    ap.add_argument("--grid", type=int, default=0, help="box, bubble: persistent grid (0: one block per tile)")
    ap.add_argument("--block", type=int, default=128, help="box, bubble: threads per list tile with --grid/--graphs")
    ap.add_argument("--dense", type=int, default=1, choices=(0, 1, 2),
                    help="box: flux op by windows for dense clusters (0 never, 1 auto, 2 always)")
    ap.add_argument("--base", type=int, default=None,
                    help="drop: base octree level (5); learned3d: 4")
    ap.add_argument("--extra", type=int, default=None,
                    help="levels above the base: drop 4, learned3d 3, bubble 4")
    ap.add_argument("--horizon", type=int, default=None, help="learned: 16, learned3d: 8 global steps")
    ap.add_argument("--train-runs", type=int, default=3)
    ap.add_argument("--train-steps", type=int, default=48)
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--rings", type=int, default=1, help="learned: face-neighbour rings added to predicted refinement")
    ap.add_argument("--every", type=int, default=8, help="learned3d: remesh cadence of the predictor run")
    ap.add_argument("--cadences", default="", help="learned3d: extra every:rings settings to compare, e.g. 4:1,2:2")
    ap.add_argument("--drop-max-dt", type=float, default=1.6e-7,
                    help="learned3d: global step cap (limits the LTS level spread; wider spreads break down at high "
                         "Mach)")
    ap.add_argument("--instances", type=int, default=64)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-tune", action="store_true", help="batched: keep the default launch (no autotuning)")
    ap.add_argument("--plain-steps", type=int, default=None,
                    help="batched: time the plain form over this many steps (default: --steps)")
    args = ap.parse_args()
    if args.steps is None:
        # the 3D shock-drop at levels 5-9 ends in a NaN state (force_break)
        # in global step 6 (at levels 4-8 the oracle breaks down in the same
        # step as the GPU, their states bitwise equal before it): time steps 3-5
        args.steps = {"drop": 3, "learned": 48, "batched": 4}.get(args.workload, 20)
    # per-workload sizes: the 3D shock-drop at levels 5-9; its learned form at
    # levels 4-7 (the feature samples of every leaf over 48 steps of three
    # runs must fit in host memory); the 2D shock-bubble 64x32 + 4 levels
    if args.base is None:
        args.base = 4 if args.workload == "learned3d" else 5
    if args.extra is None:
        args.extra = 3 if args.workload == "learned3d" else 4
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batched":
        batched(args)
        return
    if args.horizon is None:
        args.horizon = 8 if args.workload == "learned3d" else 16
    if args.workload == "learned3d":
        return learned3d(args)
    if args.workload == "learned":
        learned(args)
        return
    if args.workload == "drop":
        drop(args)
        return
    if args.workload == "bubble":
        bubble(args)
        return

    from dacti_amd import scene, workloads
    from dacti_amd.integrator import Integrator
    if args.copy_probe:
        copy_probe()

    t_build = time.perf_counter()
    case = workloads.periodic_box(level=args.level)
    tree = scene.build_tree(case)
    print(f"[bench] tree built: {tree.n} nodes in {time.perf_counter() - t_build:.1f} s", file=sys.stderr, flush=True)
    g = Integrator(case, tree)
    g.set_dense(args.dense)
    if args.graphs or args.grid:
        # a fixed grid lets the small substeps run planned on the device
        g.set_launch(block=args.block, grid=args.grid or 2048, graphs=bool(args.graphs))
    print("[bench] integrator created", file=sys.stderr, flush=True)
    for w in range(args.warmup):
        g.integrate()
        print(f"[bench] warmup step {w} done", file=sys.stderr, flush=True)
    g.set_timing(True)
    s0 = g.stats()
    t0 = time.perf_counter()
    for k in range(args.steps):
        _, fb = g.integrate()
        if fb:
            raise RuntimeError("NaN state (force_break)")
    el = time.perf_counter() - t0
    s1 = g.stats()
    print(f"[bench] {args.steps} timed steps in {el:.3f} s", file=sys.stderr, flush=True)
    updates = s1.cell_updates - s0.cell_updates
    value = updates / el
    launches = s1.flux_launches - s0.flux_launches
    kern_ms = s1.substep_kernel_ms - s0.substep_kernel_ms
    # one "launch" of the dominant work = all kernels of one level substep
    alg_bytes_per_launch = B_ALG * updates / max(launches, 1)
    avg_ms = kern_ms / max(launches, 1)
    achieved = alg_bytes_per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    # leaf-cell updates per second of substep kernel time (the roofs' rate)
    upd_rate = updates / (kern_ms * 1e-3) if kern_ms > 0 else 0.0
    pf = profile_figures()
    g.close()
    base = None
    if not args.no_cpu_baseline:
        print("[bench] timing the CPU baseline", file=sys.stderr, flush=True)
        base = cpu_baseline(case, tree)
    out = {
        "metric": "cell-updates/s (HLLC+MUSCL FP64) & %HBM roofline; learned/batched vs plain",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"3D Euler periodic box, uniform octree level {args.level} "
                               f"({(1 << args.level) ** 3} leaves), LTS on, no refinement",
                   "leaves": (1 << args.level) ** 3, "model": "euler<3>", "limiter": "minmod",
                   "max_cfl": case.max_cfl, "global_dt": s1.global_dt, "max_acti_level": s1.max_acti_level,
                   "cell_updates_timed": updates, "substeps_timed": launches,
                   "launch": {"block": args.block if (args.graphs or args.grid) else 256,
                              "grid": (args.grid or 2048) if (args.graphs or args.grid) else 0,
                              "graphs": bool(args.graphs)}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": pf["bytes"] * updates / max(launches, 1) if pf["bytes"] else None,
                     "kernel": "one level substep (list selection, flux, update, sweep, LTS apply, histogram)",
                     "bytes_per_update": B_ALG, "measured_bytes_per_update": pf["bytes"],
                     "avg_launch_ms": avg_ms,
                     # secondary roofs of the same substep kernels: FP64 arithmetic and VALU issue
                     "fp64": ({"achieved_tflops": pf["flops"] * upd_rate / 1e12, "peak_tflops": FP64_PEAK_TFLOPS,
                               "frac": pf["flops"] * upd_rate / 1e12 / FP64_PEAK_TFLOPS,
                               "flops_per_update": pf["flops"]} if pf["flops"] else None),
                     "valu_issue": ({"achieved_tlane_instr": pf["valu"] * upd_rate / 1e12,
                                     "peak_tlane_instr": VALU_PEAK_TLANE,
                                     "frac": pf["valu"] * upd_rate / 1e12 / VALU_PEAK_TLANE,
                                     "instr_per_update": pf["valu"]} if pf["valu"] else None),
                     "figures_from": "profiles/leaf_update_euler3d_256_{hbm_bytes,valu}.txt"},
        "cpu_baseline": base,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
